"""The Lua-facing API of the spectra path, mirrored in Python with the same
names, argument meaning and error behaviour (LuaInterface/ssp.cpp:572-700
NewOperator, :927 NewWavefunction, :1906-3089 CreateSpectra, :4039-4272
Eigensystem; LuaInterface/Operator.cpp:4775 and Wavefunction.cpp:200-398 for
`H*psi`, `psi*H*psi`). Every numerical step runs in libqs on the GPU; this
module only decides bases and marshals data.

Conventions kept from the reference:
* raw operator terms are written products, positive = create, negative (or -0.0)
  = annihilate; `{i, -j, v}` is v c+_i c_j;
* Restrictions are `[NF, NB, ["mask", min, max], ...]`;
* options are lists of pairs `[["Emin", -10], ["NE", 2000], ...]` (or a dict);
  an unknown option name is an error, as LuaCheckOptions makes it;
* CreateSpectra returns (Spectra, [Tri tables]) where each table holds
  a, b (shifted by -E0 on a[1..N]), mu = 0 and type = "Tri".
"""
from __future__ import annotations

import math
import warnings
from typing import Iterable

import numpy as np
import torch

from . import det as D
from . import operators as O
from ._lib import QSError, default_context
from .engine import (Basis, Batch, Hop, Op, block_cf_spectrum, broaden, cf_spectrum, dense_lowest, dense_poles,
                     pole_spectrum, tridiag_extremes)

SQRT_EPS = math.sqrt(np.finfo(float).eps)


# ---------------------------------------------------------------------------
# options
# ---------------------------------------------------------------------------
def _options(opts, allowed: dict, aliases: dict | None = None, unsupported: dict | None = None,
             given: set | None = None) -> dict:
    """Parse a list of {name, value} pairs against `allowed` (name -> default),
    as LuaCheckOptions does: an unknown name is an error; aliases map the
    reference's second spelling of a name onto the first (the later pair wins,
    both write one setting); names in `unsupported` (name -> reason) are
    options of the reference outside this engine and raise with the reason.
    `given` receives the names the caller set."""
    out = dict(allowed)
    if opts is None:
        return out
    aliases = aliases or {}
    items = opts.items() if isinstance(opts, dict) else opts
    for item in items:
        name, value = item[0], item[1]
        if unsupported and name in unsupported:
            raise QSError(f"option {name!r} is not supported: {unsupported[name]}")
        name = aliases.get(name, name)
        if name not in allowed:
            raise QSError(f"unknown option {name!r}; possible: {', '.join(list(allowed) + list(aliases))}")
        out[name] = value
        if given is not None:
            given.add(name)
    return out


def _flag_untruncated(o, given, where):
    """The reference drops determinants whose amplitude falls below Epsilon
    (ComplexWaveFunctionTruncateBasis, Core/ComplexWaveFunctionArithmetic.cpp:1073);
    this engine keeps every determinant of the space, i.e. runs at Epsilon = 0,
    the limit the reference's results approach. A caller's Epsilon is accepted
    and flagged."""
    if "Epsilon" in given and float(o["Epsilon"]) > 0:
        warnings.warn(f"{where}: Epsilon = {o['Epsilon']} is not applied; no determinant is truncated "
                      "(the engine works at Epsilon = 0)", stacklevel=3)


DBL_EPSILON = float(np.finfo(float).eps)
DBL_MIN = float(np.finfo(float).tiny)


def hermitian_class(H: "Operator", zero: float) -> int:
    """QOperatorHermitian (Core/OperatorArithmetic.cpp:6517-6568): the normal-ordered
    H^+ - H chopped at 100 DBL_MIN is empty -> 1 (Hermitian); empty after a chop
    at `zero` -> 2 (Hermitian up to zero); otherwise 0."""
    d = O.normal_order(list(dagger_terms(H.terms).items()) + [(c, -v) for c, v in H.terms.items()],
                       100 * DBL_MIN)
    if not d:
        return 1
    if all(abs(v) <= zero for v in d.values()):
        return 2
    return 0


def hermitian_part(H: "Operator", zero: float) -> "Operator":
    """The repair LuaEigensystem applies to an almost-Hermitian operator
    (LuaInterface/ssp.cpp:4164-4205): H <- normal order of (H + H^+)/2 chopped at
    DBL_EPSILON zero, twice."""
    eps = DBL_EPSILON * zero

    def sym(terms):
        t = O.normal_order(list(terms.items()) + list(dagger_terms(terms).items()))
        return {c: 0.5 * v for c, v in t.items() if abs(0.5 * v) > eps}

    return Operator(H.NF, sym(sym(H.terms)), restrictions=H.Restrictions, name=H.Name)


def _restrictions(R, nf):
    """[NF, NB, [mask, min, max], ...] -> [(mask, min, max), ...]."""
    if R is None:
        return []
    if R[0] != nf:
        raise QSError(f"restrictions are for {R[0]} fermions, the operator has {nf}")
    if R[1] != 0:
        raise QSError("bosons are not supported")
    return [(str(m), int(a), int(b)) for m, a, b in R[2:]]


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------
class Operator:
    """A many-body operator: ladder terms (codes, o[0] applied first) -> value."""

    def __init__(self, nf: int, terms: dict | Iterable = (), restrictions=None, name: str = ""):
        self.NF = nf
        self.NB = 0
        self.terms: dict = {}
        for c, v in (terms.items() if isinstance(terms, dict) else terms):
            c = tuple(int(x) for x in c)
            self.terms[c] = self.terms.get(c, 0) + complex(v)
        self.Restrictions = restrictions
        self.Name = name
        self._op = None

    # -- algebra ------------------------------------------------------------
    def _compatible(self, other):
        if other.NF != self.NF:
            raise QSError(f"operators on {self.NF} and {other.NF} fermions cannot be combined")

    def __add__(self, other):
        if isinstance(other, (int, float, complex)):
            other = Operator(self.NF, [((), other)])
        self._compatible(other)
        out = Operator(self.NF, self.terms)
        for c, v in other.terms.items():
            out.terms[c] = out.terms.get(c, 0) + v
        return out

    __radd__ = __add__

    def __neg__(self):
        return Operator(self.NF, {c: -v for c, v in self.terms.items()})

    def __sub__(self, other):
        return self + (-other if isinstance(other, Operator) else -complex(other))

    def __rsub__(self, other):
        return (-self) + other

    def __mul__(self, other):
        if isinstance(other, (int, float, complex)):
            return Operator(self.NF, {c: v * other for c, v in self.terms.items()})
        if isinstance(other, Operator):
            self._compatible(other)
            out = {}
            for ca, va in self.terms.items():
                for cb, vb in other.terms.items():
                    c = cb + ca  # other acts first
                    out[c] = out.get(c, 0) + va * vb
            return Operator(self.NF, out)
        if isinstance(other, Wavefunction):
            return operator_psi(self, other)
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, (int, float, complex)):
            return self * other
        return NotImplemented

    def __truediv__(self, other):
        return self * (1.0 / other)

    def ConjugateTranspose(self):
        return Operator(self.NF, dagger_terms(self.terms))

    def Chop(self, eps=1e-14):
        return Operator(self.NF, {c: v for c, v in self.terms.items() if abs(v) > eps})

    @property
    def NTerms(self):
        return len(self.terms)

    def term_list(self):
        return list(self.terms.items())

    def device(self, ctx=None):
        if self._op is None:
            self._op = Op(self.NF, self.term_list(), ctx or default_context())
        return self._op

    def particle_change(self):
        """Set of particle-number changes of the terms."""
        return {sum(1 if x & O.CR else -1 for x in c) for c in self.terms}

    def __repr__(self):
        return f"Operator(NF={self.NF}, terms={len(self.terms)})"


def dagger_terms(terms: dict) -> dict:
    return {tuple(x ^ O.CR for x in reversed(c)): complex(v).conjugate() for c, v in terms.items()}


def _shell_args(args):
    return list(args[0]), list(args[1])


def DeterminantString(nf: int, *index_lists) -> str:
    """A determinant string of nf modes with a 1 at every mode named in the
    index lists (LuaDeterminantString, LuaInterface/ssp.cpp:5409): the form
    restriction masks and NewWavefunction take. Indices outside [0, nf) are
    skipped with a warning, repeated ones count once."""
    if not index_lists:
        raise QSError("DeterminantString called with 1 arguments. Number of Fermions and one or several Lists "
                      "of Indices expected.")
    on = set()
    for lst in index_lists:
        for i in np.atleast_1d(lst):
            i = int(i)
            if 0 <= i < nf:
                on.add(i)
            else:
                warnings.warn(f"Index {i} is not within [0,NF-1]=[0,{nf - 1}].")
    return "".join("1" if i in on else "0" for i in range(int(nf)))


def IndexToBitMask(nf: int, indices) -> str:
    """The mask string of nf modes with the given modes set ("IndexToBitMask",
    LuaIndexToBitMask, LuaInterface/ssp.cpp:6038-6062): a negative nf, or an
    index below zero or at or above nf, is an error."""
    nf = int(nf)
    if nf < 0:
        raise QSError("IndexToBitmask called with a negative value for NF")
    idx = [int(i) for i in np.atleast_1d(indices)] if np.size(indices) else []
    for pos, i in enumerate(idx, 1):
        if i < 0:
            raise QSError(f"Index at position {pos} smaller than zero ({i}).")
        if i >= nf:
            raise QSError(f"Index at position {pos} larger or equal than NF ({nf},{i}).")
    on = set(idx)
    return "".join("1" if i in on else "0" for i in range(nf))


def BitMaskToIndex(mask: str):
    """(NF, [set modes]) of a mask string: NF counts the '0' and '1' characters,
    every other character is skipped (LuaBitMaskToIndex, LuaInterface/ssp.cpp:6017;
    Core/Utilities.cpp:5-31)."""
    bits = [c for c in str(mask) if c in "01"]
    return len(bits), [i for i, c in enumerate(bits) if c == "1"]


def PotentialExpandedOnClm(group: str, l: int, energies):
    """A crystal-field potential on renormalized spherical harmonics,
    [[k, m, A_km], ...], whose eigenvalues on an l shell are the given energies of
    the irreducible representations, in the reference's order and orientation
    (LuaPotentialExpandedOnClm, LuaInterface/ssp.cpp:908; the input to
    NewOperator("CF", ...)). Groups: C1 and Ci (one energy, any l), Oh, O and Td
    (l <= 1 one energy; l = 2 {e_g, t_2g}), D4h, C4v, D4 and D2d (l = 1 {a, e};
    l = 2 {a1 (z^2), b1 (x^2-y^2), b2 (xy), e (xz, yz)}), D2h (l = 2 {a_g x^2-y^2,
    a_g z^2, their mixing, b1g xy, b2g xz, b3g yz}) and D3d (l = 2 {a1g, eg(1),
    eg(2), their mixing}); see crystal_field.py for the construction."""
    from .crystal_field import potential_expanded_on_clm
    try:
        return potential_expanded_on_clm(str(group), int(l), energies)
    except KeyError:
        raise QSError(f"Potential of \"{group}\" symmetry for angular momentum {l} not found") from None
    except ValueError as e:
        raise QSError(f"PotentialExpandedOnClm(\"{group}\", {l}, ...): {e}") from None


def NewOperator(*args, **kw):
    """NewOperator(NF, NB, {{i, -j, ..., v}, ...}) or a named operator:
    "Cr"/"An" (NF, i), "Number" (NF, idxA, idxB, values), "U" (NF, up, dn, F) or
    (NF, upA, dnA, upB, dnB, F, G), "ldots", "Sx"/"Sy"/"Sz"/"Splus"/"Smin"/"Ssqr" and the
    same for L and J (NF, up, dn),
    "CF" (NF, up, dn, Akm) or (NF, upA, dnA, upB, dnB, Akm)."""
    if not args:
        raise QSError("NewOperator called with 0 arguments, expect at least one")
    if isinstance(args[0], (int, np.integer)):
        if len(args) not in (3, 4):
            raise QSError(f"NewOperator called with {len(args)} arguments, expected 3 or 4")
        nf, nb, table = int(args[0]), int(args[1]), args[2]
        if nb != 0:
            raise QSError("bosons are not supported")
        opts = _options(args[3] if len(args) == 4 else None, {"Restrictions": None, "Name": ""})
        out = {}
        for pos, t in enumerate(table, 1):
            if len(t) == 0:
                raise QSError(f"NewOperator called an empty table as operator at position {pos}")
            *ops, v = t
            codes = []
            for j, o in enumerate(ops, 1):
                if abs(o) > nf - 1 or abs(abs(o) - round(abs(o))) > 10 * np.finfo(float).eps:
                    raise QSError(f"Operator {pos} at position {j} has invalid index {o}")
                idx = int(round(abs(o)))
                codes.append(idx if math.copysign(1.0, o) < 0 else (O.CR | idx))
            c = tuple(reversed(codes))  # the rightmost written factor acts first
            out[c] = out.get(c, 0) + complex(v)
        return Operator(nf, out, restrictions=opts["Restrictions"], name=opts["Name"])
    name, rest = args[0], args[1:]
    nf = int(rest[0])
    if name in ("Cr", "An"):
        i = int(rest[1])
        return Operator(nf, [((O.cr(i) if name == "Cr" else O.an(i),), 1.0)])
    if name == "Number":
        a, b = list(rest[1]), list(rest[2])
        vals = list(rest[3]) if len(rest) > 3 else [1.0] * len(a)
        return Operator(nf, [O.written([O.cr(i), O.an(j)], v) for i, j, v in zip(a, b, vals)])
    if name in ("U", "Coulomb"):
        if len(rest) == 4:
            up, dn = _shell_args(rest[1:3])
            return Operator(nf, O.coulomb(up, dn, list(rest[3])))
        upa, dna, upb, dnb = (list(x) for x in rest[1:5])
        return Operator(nf, O.coulomb_two_shells(upa, dna, upb, dnb, list(rest[5]), list(rest[6])))
    if name in ("ldots", "Lx", "Ly", "Lz", "Sx", "Sy", "Sz", "Jx", "Jy", "Jz", "Splus", "Smin", "Lplus", "Lmin",
                "Jplus", "Jmin"):
        up, dn = _shell_args(rest[1:3])
        return Operator(nf, O.shell_operator(up, dn, name))
    if name in ("Ssqr", "Lsqr", "Jsqr"):
        # the square of the vector operator as products of its one-body components
        up, dn = _shell_args(rest[1:3])
        k = name[0]
        comps = [Operator(nf, O.shell_operator(up, dn, k + c)) for c in "xyz"]
        return comps[0] * comps[0] + comps[1] * comps[1] + comps[2] * comps[2]
    if name == "CF":
        if len(rest) == 4:
            up, dn = _shell_args(rest[1:3])
            return Operator(nf, O.dipole_p_to_d(up, dn, up, dn, [tuple(x) for x in rest[3]]))
        upa, dna, upb, dnb = (list(x) for x in rest[1:5])
        return Operator(nf, O.dipole_p_to_d(upa, dna, upb, dnb, [tuple(x) for x in rest[5]]))
    raise QSError(f"unknown operator name {name!r}")


# ---------------------------------------------------------------------------
# wavefunctions
# ---------------------------------------------------------------------------
class Wavefunction:
    """A wavefunction: sorted determinants and complex coefficients. It may be
    host-resident (words, coef) or device-resident (Basis, torch vector)."""

    def __init__(self, nf, words=None, coef=None, basis: Basis | None = None, vec: torch.Tensor | None = None):
        self.NF = nf
        self.NB = 0
        self._words = words
        self._coef = coef
        self._basis = basis
        self._vec = vec

    # -- storage ------------------------------------------------------------
    def host(self):
        if self._words is None:
            w = self._basis.words()
            c = self._vec.detach().cpu().numpy()
            keep = c != 0
            self._words, self._coef = w[keep], c[keep]
        return self._words, self._coef

    def device(self, ctx=None):
        if self._basis is None:
            ctx = ctx or default_context()
            w, c = self.host()
            self._basis = Basis.from_words(self.NF, w, ctx)
            # from_words sorts: place the coefficients by lookup
            idx = self._basis.index(w)
            v = np.zeros(self._basis.n, dtype=np.complex128)
            v[idx] = c
            self._vec = torch.from_numpy(v).to(torch.device("cuda", ctx.device))
        return self._basis, self._vec

    @property
    def N(self):
        if self._words is not None:
            return len(self._words)
        return int(torch.count_nonzero(self._vec).item())

    def particle_numbers(self):
        w, _ = self.host()
        if len(w) == 0:
            return set()
        return {int(x) for x in np.unique(np.bitwise_count(w).sum(axis=1))}

    # -- algebra ------------------------------------------------------------
    def _combine(self, other, fa, fb):
        wa, ca = self.host()
        wb, cb = other.host()
        keys = {}
        for w, c in ((wa, ca * fa), (wb, cb * fb)):
            for row, v in zip(map(tuple, w.tolist()), c):
                keys[row] = keys.get(row, 0) + v
        rows = sorted(keys)
        words = np.array(rows, dtype=np.uint64).reshape(-1, D.nwords(self.NF))
        return Wavefunction(self.NF, words, np.array([keys[r] for r in rows], dtype=np.complex128))

    def __add__(self, other):
        return self._combine(other, 1.0, 1.0)

    def __sub__(self, other):
        return self._combine(other, 1.0, -1.0)

    def __mul__(self, other):
        if isinstance(other, (int, float, complex)):
            if self._words is not None:
                return Wavefunction(self.NF, self._words, self._coef * other)
            return Wavefunction(self.NF, basis=self._basis, vec=self._vec * other)
        if isinstance(other, Operator):
            # psi * H is the bra <psi|H, represented by H^+ |psi> (LuaWavefunctionMul)
            return operator_psi(Operator(other.NF, dagger_terms(other.terms), other.Restrictions), self)
        if isinstance(other, Wavefunction):
            return dot(self, other)
        return NotImplemented

    __rmul__ = lambda self, other: self * other if isinstance(other, (int, float, complex)) else NotImplemented

    def __truediv__(self, other):
        return self * (1.0 / other)

    def Normalize(self):
        n = math.sqrt(abs(dot(self, self)))
        if self._words is not None:
            self._coef = self._coef / n
        else:
            self._vec = self._vec / n
        return self

    def __eq__(self, other):
        d = self - other
        _, c = d.host()
        return bool(np.all(np.abs(c) <= 100 * np.finfo(float).eps * max(1.0, float(np.abs(self.host()[1]).max(initial=0)))))

    def __repr__(self):
        return f"Wavefunction(NF={self.NF}, N={self.N})"


def NewWavefunction(nf, nb, table, options=None):
    """NewWavefunction(NF, NB, {{"0101", coef}, ...}) (LuaInterface/ssp.cpp:927)."""
    if nb != 0:
        raise QSError("NewWavefunction with NB > 0 is not supported (the reference rejects it too)")
    keys = {}
    for s, c in table:
        w = D.string_to_words(str(s), nf)
        keys[w] = keys.get(w, 0) + complex(c)
    rows = sorted(keys)
    words = np.array(rows, dtype=np.uint64).reshape(-1, D.nwords(nf))
    return Wavefunction(nf, words, np.array([keys[r] for r in rows], dtype=np.complex128))


def Braket(*args):
    """Braket(a, b) = <a|b> or Braket(a, O, b) = <a|O|b> (the "Braket" global,
    LuaInterface/ssp.cpp:6131). a and b may each be a wavefunction or a list of
    them: a list on one side gives a list, lists on both sides a table
    [i][j] = <a_i|(O)|b_j>. O|b_j> is formed once per b_j."""
    if len(args) == 2:
        a, b = args
        op = None
    elif len(args) == 3:
        a, op, b = args
    else:
        raise QSError(f"Braket called with {len(args)} arguments, expected 2 or 3")
    a_list, b_list = isinstance(a, (list, tuple)), isinstance(b, (list, tuple))
    bs = list(b) if b_list else [b]
    if op is not None:
        bs = [operator_psi(op, x) for x in bs]
    rows = [[dot(x, y) for y in bs] for x in (list(a) if a_list else [a])]
    if a_list and b_list:
        return rows
    if a_list:
        return [r[0] for r in rows]
    return rows[0] if b_list else rows[0][0]


def BraketDiagonal(*args):
    """Braket with lists on both sides paired element by element: the list
    [<a_i|(O)|b_i>] ("BraketDiagonal", LuaInterface/ssp.cpp:6132); one
    wavefunction on either side gives the same as Braket."""
    a, b = args[0], args[-1]
    if isinstance(a, (list, tuple)) and isinstance(b, (list, tuple)):
        if len(a) != len(b):
            raise QSError(f"BraketDiagonal needs lists of equal length, got {len(a)} and {len(b)}")
        return [Braket(x, *args[1:-1], y) for x, y in zip(a, b)]
    return Braket(*args)


def dot(a: Wavefunction, b: Wavefunction):
    """<a|b>, a conjugated; real when |Im| < 10(eps|Re| + DBL_MIN) (LuaWavefunctionMul)."""
    if a._basis is not None and a._basis is b._basis:
        va, vb = a._vec, b._vec
        r = complex(torch.vdot(va, vb).item())
    else:
        wa, ca = a.host()
        wb, cb = b.host()
        ia = {tuple(x): i for i, x in enumerate(wa.tolist())}
        r = 0j
        for j, x in enumerate(wb.tolist()):
            i = ia.get(tuple(x))
            if i is not None:
                r += np.conj(ca[i]) * cb[j]
    if abs(r.imag) < 10 * (np.finfo(float).eps * abs(r.real) + np.finfo(float).tiny):
        return r.real
    return r


# ---------------------------------------------------------------------------
# bases
# ---------------------------------------------------------------------------
def _sector(nf, restrictions, npart, ctx, max_dim=50_000_000) -> Basis:
    """Sector of `npart` particles inside the restrictions."""
    restr = list(restrictions) + [("1" * nf, npart, npart)]
    n = Basis.count(nf, restr)
    if n > max_dim:
        raise QSError(f"the {npart}-particle space has {n} determinants; give Restrictions to reduce it")
    return Basis.enumerate(nf, restr, ctx)


def conserved_groups(H: Operator):
    """Finest partition of the modes found by merging, term by term, the groups
    whose occupation a term changes; the occupation of every group is then
    conserved by H. The reference grows the space of Eigensystem and
    CreateSpectra as the closure of H*psi (Core/BlockLanczos.cpp:5849-5858,
    LuaInterface/ssp.cpp:2265-2316); fixing these occupations reproduces that
    closure's sector without iterating."""
    nf = H.NF
    parent = list(range(nf))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    changed = True
    while changed:
        changed = False
        for c in H.terms:
            net = {}
            for x in c:
                g = find(x & 0x7FFF)
                net[g] = net.get(g, 0) + (1 if x & O.CR else -1)
            moving = [g for g, d in net.items() if d != 0]
            for g in moving[1:]:
                a, b = find(moving[0]), find(g)
                if a != b:
                    parent[b] = a
                    changed = True
    groups = {}
    for m in range(nf):
        groups.setdefault(find(m), []).append(m)
    return sorted(groups.values())


def _group_mask(nf, modes):
    s = ["0"] * nf
    for m in modes:
        s[m] = "1"
    return "".join(s)


def _sector_of_counts(nf, groups, counts, restrictions, ctx, max_dim=50_000_000) -> Basis:
    # groups held empty (or full) form one restriction together: count 0 on the
    # union (count = its size); operators leave many modes untouched
    empty = [m for g, c in zip(groups, counts) if c == 0 for m in g]
    full = [m for g, c in zip(groups, counts) if c == len(g) and c > 0 for m in g]
    rest = [(g, c) for g, c in zip(groups, counts) if 0 < c < len(g)]
    restr = list(restrictions) + [(_group_mask(nf, g), c, c) for g, c in rest]
    if empty:
        restr.append((_group_mask(nf, empty), 0, 0))
    if full:
        restr.append((_group_mask(nf, full), len(full), len(full)))
    n = Basis.count(nf, restr)
    if n > max_dim:
        raise QSError(f"the reachable space has {n} determinants; give Restrictions to reduce it")
    return Basis.enumerate(nf, restr, ctx)


def _counts_of_words(nf, groups, words):
    """Occupation of every group for each determinant (host words)."""
    ints = [D.words_to_int(w) >> (64 * D.nwords(nf) - nf) for w in words]
    masks = [int(_group_mask(nf, g), 2) for g in groups]
    return {tuple(bin(x & m).count("1") for m in masks) for x in ints}


def operator_psi(op: Operator, psi: Wavefunction, restrictions=None) -> Wavefunction:
    """O|psi> (OperatorPsi / OperatorPsiRestricted, Core/OperatorWaveFunction.cpp:5573-5625):
    the output holds the determinants the operator's terms reach from psi's
    determinants (the reference's push form grows exactly that set), inside the
    operator's or the given restrictions. When psi is large and fills much of
    its particle sector, the sector itself (cheaper to enumerate than the
    image) is the output basis; the coefficients are the same, the extra
    determinants hold zeros."""
    ctx = default_context()
    if op.NF != psi.NF:
        raise QSError("operator and wavefunction have different numbers of fermions")
    nf = op.NF
    R = _restrictions(restrictions if restrictions is not None else op.Restrictions, nf)
    cols, x = psi.device(ctx)
    rows = None
    dn = op.particle_change()
    if cols.n >= _SECTOR_ROUTE_MIN and len(dn) == 1:
        npsi = psi.particle_numbers()
        if len(npsi) == 1:
            n_out = npsi.pop() + dn.pop()
            restr = list(R) + [("1" * nf, n_out, n_out)]
            nsec = Basis.count(nf, restr)
            if 0 <= nsec <= min(50_000_000, _SECTOR_ROUTE_RATIO * cols.n):
                rows = Basis.enumerate(nf, restr, ctx)
    if rows is None:
        rows = Basis.image(op.device(ctx), cols, R or None)
    if rows.n == 0:
        return Wavefunction(nf, np.zeros((0, D.nwords(nf)), dtype=np.uint64), np.zeros(0, dtype=np.complex128))
    hop = Hop(op.device(ctx), rows, cols, mode="matfree")
    y = hop.apply(x)
    return Wavefunction(nf, basis=rows, vec=y)


# psi of at least this many determinants whose particle sector (after the
# operator) holds at most this many times as many use the sector as output basis
_SECTOR_ROUTE_MIN = 1 << 16
_SECTOR_ROUTE_RATIO = 64


def _expect(op: Operator, psi: Wavefunction) -> complex:
    """<psi|O|psi> over psi's own support (PsiHermitianOperatorPsi)."""
    ctx = default_context()
    b, x = psi.device(ctx)
    hop = Hop(op.device(ctx), b, b, mode="matfree")
    y = hop.apply(x)
    return complex(torch.vdot(x, y).item())


# ---------------------------------------------------------------------------
# Eigensystem
# ---------------------------------------------------------------------------
_EIG_OPTIONS = {"Restrictions": None, "Zero": 100 * SQRT_EPS, "Epsilon": SQRT_EPS / 100, "SingularValue": SQRT_EPS,
                "ExpandBasis": True, "NKrylovMax": 500, "NKrylovStart": 100, "NKrylovStep": 10, "NKrylovMin": 50,
                "DenseBorder": 1513, "NBitsKey": 16, "ShiftScheme": "Variance", "Seed": 1, "MaxRestarts": 200}
_EIG_ALIASES = {"restrictions": "Restrictions", "zero": "Zero", "epsilon": "Epsilon"}


def _start_sectors(nf, groups, start, limit=256):
    """Occupation vectors of the conserved mode groups that determinants inside
    the start restrictions can have: per group the counts the start allows,
    combined and kept when some determinant has all of them at once."""
    per = []
    for g in groups:
        m = _group_mask(nf, g)
        per.append([c for c in range(len(g) + 1) if Basis.count(nf, start + [(m, c, c)]) > 0])
    combos = [()]
    for g, cs in zip(groups, per):
        combos = [t + (c,) for t in combos for c in cs]
        if len(combos) > 16 * limit:
            raise QSError(f"StartRestrictions span more than {limit} sectors of the modes H conserves")
    masks = [_group_mask(nf, g) for g in groups]
    out = [c for c in combos
           if Basis.count(nf, start + [(m, k, k) for m, k in zip(masks, c)]) > 0] if len(combos) > 1 else combos
    if not out or not per or any(not cs for cs in per):
        raise QSError("no determinant satisfies the StartRestrictions")
    if len(out) > limit:
        raise QSError(f"StartRestrictions span {len(out)} sectors of the modes H conserves; at most {limit}")
    return out


def _lowest(H: Operator, basis: Basis, k: int, o: dict, ctx):
    """The k lowest eigenpairs of H on one basis: the determinant matrix diagonalized
    up to DenseBorder determinants (BlockLanczosGroundStateConserveBasisDense,
    Core/BlockLanczos.cpp:370-504, 5713), thick-restart block Lanczos above."""
    if basis.n <= max(int(o["DenseBorder"]), k):
        Hd = Hop(H.device(ctx), basis, mode="matfree").dense()
        E, V = dense_lowest(Hd, k, ctx)
        return E[:k], [V[i].contiguous() for i in range(k)]
    hop = Hop(H.device(ctx), basis, mode="auto")
    E, vecs, _, _ = hop.ground_states(k, tol=float(o["Zero"]), max_krylov=int(o["NKrylovMax"]),
                                      max_restarts=int(o["MaxRestarts"]), seed=int(o["Seed"]),
                                      krylov_start=int(o["NKrylovStart"]), krylov_step=int(o["NKrylovStep"]),
                                      cycle_end=True)
    return E, [vecs[i].clone() for i in range(k)]


def Eigensystem(H: Operator, StartRestrictions, Npsi: int, options=None):
    """Lowest Npsi eigenstates (LuaEigensystem, LuaInterface/ssp.cpp:4039-4272).

    * Options as the reference spells them (both spellings of Restrictions,
      Zero and Epsilon). NKrylovMin is accepted and, as in the reference, not
      used; NBitsKey sizes the reference's hash tables, which this engine does
      not have; ShiftScheme ("Automatic", "Variance", "None") shifts the
      reference's Krylov iteration and changes its convergence path only, the
      thick-restart solver here needs no shift; Epsilon is flagged when set
      (no truncation here); SingularValue is accepted (the block solver keeps
      its start block full rank by itself).
    * H is checked for Hermiticity as the reference does: not Hermitian up to
      Zero raises; Hermitian up to Zero is repaired into its Hermitian part.
    * The space: with ExpandBasis (default) every sector of the mode groups H
      conserves that the StartRestrictions reach (the closure of H*psi the
      reference grows, Core/BlockLanczos.cpp:5849-5858), inside the
      Restrictions; with ExpandBasis = false the determinants of the
      StartRestrictions themselves (BlockLanczosGroundStateConserveBasis).
      The Npsi lowest states of every sector are found and the Npsi lowest of
      all are returned, ascending in energy."""
    given: set = set()
    o = _options(options, _EIG_OPTIONS, _EIG_ALIASES, given=given)
    _flag_untruncated(o, given, "Eigensystem")
    if str(o["ShiftScheme"]) not in ("Automatic", "Variance", "None"):
        raise QSError(f'Eigensystem: unknown ShiftScheme "{o["ShiftScheme"]}", possible are "Automatic", '
                      '"Variance" and "None"')
    nf = H.NF
    Npsi = int(Npsi)
    if Npsi < 1:
        raise QSError("Eigensystem needs Npsi >= 1")
    if H.particle_change() - {0}:
        raise QSError("Eigensystem needs a particle-number conserving Hamiltonian")
    herm = hermitian_class(H, float(o["Zero"]))
    if herm == 0:
        raise QSError("Eigensystem called with an operator that is not Hermitian")
    if herm == 2:
        H = hermitian_part(H, float(o["Zero"]))
    start = _restrictions(StartRestrictions, nf)
    R = _restrictions(o["Restrictions"], nf)
    ctx = default_context()
    if not o["ExpandBasis"]:
        bases = [Basis.enumerate(nf, start + R, ctx)]
    else:
        groups = conserved_groups(H)
        bases = [_sector_of_counts(nf, groups, c, R, ctx) for c in _start_sectors(nf, groups, start)]
    found = []
    total = 0
    for basis in bases:
        total += basis.n
        k = min(Npsi, basis.n)
        if k == 0:
            continue
        E, vecs = _lowest(H, basis, k, o, ctx)
        found += [(float(E[i]), len(found), Wavefunction(nf, basis=basis, vec=vecs[i])) for i in range(k)]
    if len(found) < Npsi:
        raise QSError(f"the space has {total} determinants, fewer than Npsi = {Npsi}")
    found.sort(key=lambda t: (t[0], t[1]))
    return [w for _, _, w in found[:Npsi]]


# ---------------------------------------------------------------------------
# CreateSpectra
# ---------------------------------------------------------------------------
class TriResponse(dict):
    """A response function in tridiagonal form: {1: a, 2: b, mu, type="Tri", name}
    callable as G(omega, gamma) (EvaluateTriDiagonalMatrix, BasicMath_GreenFunctionArithmetic.cpp:4621)."""

    def __init__(self, a, b, mu=0.0, name=""):
        super().__init__(a=np.asarray(a, dtype=float), b=np.asarray(b, dtype=float), mu=mu, type="Tri", name=name)

    def __call__(self, omega, gamma):
        g = cf_spectrum(self["a"], self["b"], 0.0, float(omega), float(omega) + 1.0, 1, float(gamma))
        return complex(g[0, 0])


class PolesResponse(dict):
    """A response function as a list of poles: {a: [a0, E_1, ..., E_N], b: [w_1, ..., w_N],
    mu, type="ListOfPoles", name} (LuaPushTable_ListOfPoles,
    LuaInterface/lua_extensions/TableToArray.cpp:3640-3669), callable as
    G(omega, gamma) = a0 + sum_k w_k / (omega - E_k + i gamma/2)."""

    def __init__(self, a, b, mu=0.0, name=""):
        super().__init__(a=np.asarray(a, dtype=float), b=np.asarray(b, dtype=float), mu=mu, type="ListOfPoles",
                         name=name)

    def __call__(self, omega, gamma):
        P = np.sqrt(self["b"]).astype(np.complex128)[:, None]
        g = pole_spectrum(self["a"][1:], P, 0.0, float(omega), float(omega) + 1.0, 1, float(gamma))
        return complex(g[0, 0]) + float(self["a"][0])


class Spectra:
    """Spectra on the grid w_i = Emin + i (Emax-Emin)/NE, i = 0..NE (SpectraTypeArray)."""

    def __init__(self, emin, emax, ne, gamma, G, eminpole=None, emaxpole=None):
        self.Emin, self.Emax, self.NE, self.Gamma = emin, emax, ne, gamma
        self.G = np.asarray(G)  # (nspec, NE+1) complex
        self.EminPole, self.EmaxPole = eminpole, emaxpole

    @property
    def energies(self):
        return self.Emin + np.arange(self.NE + 1) * (self.Emax - self.Emin) / self.NE

    def Sum(self, *weights):
        """Spectra.Sum(s, w_1, w_2, ...) (LuaInterface/Spectra.cpp:527-613): one
        spectrum per weight list, sum_k w_j[k] s[k], real or complex weights."""
        if not weights:
            raise QSError("Spectra.Sum needs at least one list of weights")
        G = self.G.reshape(-1, self.NE + 1)
        out = []
        for j, w in enumerate(weights, 1):
            w = np.asarray(w, dtype=np.complex128).ravel()
            if w.size != G.shape[0]:
                raise QSError(f"Spectra.Sum: weight list {j} has {w.size} entries for {G.shape[0]} spectra")
            out.append(w @ G)
        return Spectra(self.Emin, self.Emax, self.NE, self.Gamma, out, self.EminPole, self.EmaxPole)

    # arithmetic (LuaInterface/Spectra.cpp:969-1752): a number adds to the real
    # parts; spectra combine point by point
    def _other(self, x):
        if isinstance(x, Spectra):
            if x.G.shape != self.G.shape or (x.Emin, x.Emax, x.NE) != (self.Emin, self.Emax, self.NE):
                raise QSError("spectra on different grids or of different counts cannot be combined")
            return x.G
        return x

    def _new(self, G):
        return Spectra(self.Emin, self.Emax, self.NE, self.Gamma, G, self.EminPole, self.EmaxPole)

    def __add__(self, x):
        return self._new(self.G + self._other(x) if isinstance(x, Spectra) else self.G + float(np.real(x)))

    __radd__ = __add__

    def __sub__(self, x):
        return self._new(self.G - self._other(x) if isinstance(x, Spectra) else self.G - float(np.real(x)))

    def __mul__(self, x):
        return self._new(self.G * self._other(x))

    __rmul__ = __mul__

    def __truediv__(self, x):
        return self._new(self.G / self._other(x))

    def Copy(self):
        return Spectra(self.Emin, self.Emax, self.NE, self.Gamma, self.G.copy(), self.EminPole, self.EmaxPole)

    Clone = Copy

    def Shift(self, shift):
        """The spectra on an energy axis moved by `shift` (a number, or one per
        spectrum, all equal here since the spectra share one grid); the values
        stay (LuaInterface/Spectra.cpp:217-274)."""
        sh = np.atleast_1d(np.asarray(shift, dtype=float))
        if sh.size > 1 and (sh.size != self.G.reshape(-1, self.NE + 1).shape[0] or np.any(sh != sh[0])):
            raise QSError("Spectra.Shift: the spectra share one energy grid; give one shift")
        d = float(sh[0])
        pmin = None if self.EminPole is None else self.EminPole + d
        pmax = None if self.EmaxPole is None else self.EmaxPole + d
        return Spectra(self.Emin + d, self.Emax + d, self.NE, self.Gamma, self.G.copy(), pmin, pmax)

    def MultiplyByEnergy(self):
        """Every value times its energy E_j = Emin + j dE (Spectra.cpp:484-505)."""
        return self._new(self.G * self.energies)

    def Element(self, which):
        """Spectrum `which` (1-based), or the listed ones (Spectra.cpp:695-811)."""
        G = self.G.reshape(-1, self.NE + 1)
        idx = [which] if isinstance(which, (int, np.integer)) else list(which)
        if any(i < 1 or i > G.shape[0] for i in idx):
            raise QSError(f"Spectra.Element: no element {idx} in spectra array with {G.shape[0]} elements")
        return Spectra(self.Emin, self.Emax, self.NE, self.Gamma, G[[i - 1 for i in idx]].copy(), self.EminPole,
                       self.EmaxPole)

    def ToTable(self):
        """One list per spectrum of [energy, complex value] pairs (Spectra.cpp:325-335)."""
        E = self.energies
        return [[[float(e), complex(v)] for e, v in zip(E, g)] for g in self.G.reshape(-1, self.NE + 1)]

    @staticmethod
    def FromTable(table):
        """Spectra from ToTable's form (Spectra.cpp:337-480); the spectra must share one
        equidistant grid."""
        if not table or any(len(sp) < 2 for sp in table):
            raise QSError("Spectra.FromTable needs spectra of at least two points")
        E = np.array([[float(p[0]) for p in sp] for sp in table])
        if np.any(E != E[0]):
            raise QSError("Spectra.FromTable: the spectra must share one energy grid")
        ne = E.shape[1] - 1
        G = np.array([[complex(p[1]) for p in sp] for sp in table])
        return Spectra(float(E[0, 0]), float(E[0, -1]), ne, 0.0, G)

    def _widths(self, w, which):
        """A FWHM as one number, or per grid point from a table {{E_1, FWHM_1}, ...}
        (ascending energies): linear between the two entries around a point,
        the first (last) FWHM before (after) the table (LuaInterface/Spectra.cpp:37-62)."""
        if isinstance(w, (int, float, np.floating, np.integer)):
            return float(w)
        t = np.asarray(w, dtype=object)
        if t.ndim != 2 or t.shape[1] != 2 or any(isinstance(x, complex) for x in t.ravel()):
            raise QSError(f"Spectra.Broaden: the {which} width should be a number or a table "
                          "{{Energy_1, FWHM_1}, ...} of real numbers")
        t = t.astype(float)
        E, F = t[:, 0], t[:, 1]
        if np.any(np.diff(E) < 0):
            raise QSError(f"Spectra.Broaden: the energies of the {which} width table should ascend")
        x = self.Emin + np.arange(self.NE + 1) * ((self.Emax - self.Emin) / self.NE)
        k = np.searchsorted(E, x, side="left")  # table entries strictly below x
        lo = np.clip(k - 1, 0, len(E) - 1)
        hi = np.clip(k, 0, len(E) - 1)
        with np.errstate(divide="ignore", invalid="ignore"):
            mid = F[lo] + (F[hi] - F[lo]) * (x - E[lo]) / (E[hi] - E[lo])
        return np.where(k == 0, F[0], np.where(k == len(E), F[-1], mid))

    def Broaden(self, gauss, lorentz):
        """Gaussian, then Lorentzian broadening of every spectrum, in place
        (the method form spectra.Broaden(G, L), LuaInterface/Spectra.cpp:118-215;
        the function form Spectra.Broaden(s, G, L) broadens a copy: use
        s.Copy().Broaden(G, L)). Each width is a FWHM or a table (see _widths);
        a single width narrower than half the grid step leaves the spectra as
        they are, table widths are raised to half the step
        (Core/Spectral.cpp:958-1131). Computed on the GPU."""
        dE = (self.Emax - self.Emin) / self.NE
        g = self._widths(gauss, "Gaussian")
        l_ = self._widths(lorentz, "Lorentzian")
        shape = self.G.shape
        G = self.G.reshape(-1, self.NE + 1)
        G = broaden(G, dE, "G", g)
        G = broaden(G, dE, "L", l_)
        self.G = G.reshape(shape)
        return self

    def Print(self, file=None):
        """The reference's spectra table (StreamPrintSpectra, IO/StreamPrint.cpp:4-60):
        a header with the count, Emin/Emax, the pole range and dE/Gamma of every
        spectrum, then one line per energy with Re and Im of every spectrum.
        Returns the text; written to `file` when given ({{"file", name}})."""
        G = self.G.reshape(-1, self.NE + 1)
        ns = G.shape[0]
        dE = (self.Emax - self.Emin) / self.NE
        pmin = 0.0 if self.EminPole is None else self.EminPole
        pmax = 0.0 if self.EmaxPole is None else self.EmaxPole
        out = ["#Spectra: %i\n" % ns,
               "Emin______Emax     " + "".join(" %22.15E %22.15E" % (self.Emin, self.Emax) for _ in range(ns)) + "\n",
               "EminPole__EmaxPole " + "".join(" %22.15E %22.15E" % (pmin, pmax) for _ in range(ns)) + "\n",
               "dE________Gamma    " + "".join(" %22.15E %22.15E" % (dE, self.Gamma) for _ in range(ns)) + "\n"]
        head = "Energy             "
        for i in range(min(ns, 10000)):  # labels as the reference writes them (up to 9999)
            w = len(str(i))
            head += "  Re[%i]" % i + " " * (19 - w) + "Im[%i]" % i + " " * (17 - w)
        out.append(head + "\n")
        for j in range(self.NE + 1):
            row = "%19.12E" % (dE * j + self.Emin)
            for g in G:
                row += " %22.15E %22.15E" % (g[j].real, g[j].imag)
            out.append(row + "\n")
        text = "".join(out)
        if file is not None:
            with open(file, "w") as f:
                f.write(text)
        return text


def _pole_range(a, b, E0):
    n = len(b)
    lo, hi = tridiag_extremes(a[1:n + 1], b[1:n])
    return lo - E0, hi - E0


def _default_range(o, eminpole, emaxpole):
    """Emin/Emax from the pole range when not given (LuaInterface/ssp.cpp:2405-2430)."""
    emin, emax = o["Emin"], o["Emax"]
    if emin is None or emax is None:
        span = emaxpole - eminpole
        if span == 0:
            dmin, dmax = (-1.0, 1.0) if eminpole == 0 else ((2.125 * eminpole, -0.125 * eminpole) if eminpole < 0
                                                             else (-0.125 * eminpole, 2.125 * eminpole))
        else:
            dmin, dmax = eminpole - 0.125 * span, emaxpole + 0.125 * span
        emin = dmin if emin is None else emin
        emax = dmax if emax is None else emax
    return float(emin), float(emax)


def _dense_spectra(H, cb, tpsi, psiEn, o, ctx):
    """The dense method on the closure basis cb of T psi (LuaInterface/ssp.cpp:2344-2584):
    the determinant matrix of H, its eigen-decomposition, the pole weights of every
    T psi, and the spectra from the poles. Returns (Spectra, tables[, eigenstates])."""
    n, nsp, nf = cb.n, len(tpsi), H.NF
    Xh = np.zeros((nsp, n), dtype=np.complex128)
    for i, tp in enumerate(tpsi):
        w, c = tp.host()
        if len(w):
            idx = cb.index(w)
            if np.any(idx < 0):
                raise QSError("T psi has determinants outside the closure of its basis")
            Xh[i, idx] = c
    dev = torch.device("cuda", ctx.device)
    X = torch.from_numpy(Xh).to(dev)
    hop = Hop(H.device(ctx), cb, mode="matfree")
    Hd = hop.dense()
    E, P = dense_poles(Hd, X, ctx, vectors=bool(o["KrylovBasis"]))
    ne = int(o["NE"])
    gamma = float(o["Gamma"])
    if not o["Tensor"]:
        eminpole = min(E[0] - e for e in psiEn)
        emaxpole = max(E[-1] - e for e in psiEn)
        emin, emax = _default_range(o, eminpole, emaxpole)
        if gamma == 0:
            gamma = 10 * (emax - emin) / ne
        G = pole_spectrum(E, P, np.array(psiEn), emin, emax, ne, gamma, ctx=ctx)
        tables = [PolesResponse(np.r_[0.0, E - psiEn[i]], P[:, i].real ** 2 + P[:, i].imag ** 2, 0.0,
                                f"Spectrum {i}") for i in range(nsp)]
        out = [Spectra(emin, emax, ne, gamma, G, eminpole, emaxpole), tables]
    else:
        E0 = psiEn[0]
        eminpole, emaxpole = E[0] - E0, E[-1] - E0
        emin, emax = _default_range(o, eminpole, emaxpole)
        if gamma == 0:
            gamma = 10 * (emax - emin) / ne
        G = pole_spectrum(E, P, E0, emin, emax, ne, gamma, tensor=True, ctx=ctx)
        tables = {"type": "BlockListOfPoles", "E": E - E0, "blocks": np.conj(P)[:, :, None] * P[:, None, :]}
        out = [Spectra(emin, emax, ne, gamma, G.reshape(ne + 1, -1).T, eminpole, emaxpole), tables]
    if o["KrylovBasis"]:
        out.append([Wavefunction(nf, basis=cb, vec=Hd[:, k].contiguous()) for k in range(n)])
    return tuple(out)


def _block_tridiag_eigenvalues(A, B, kept, ctx):
    """Eigenvalues (ascending) of the block tridiagonal matrix of a block Lanczos
    run: diagonal blocks A[l] (kept[l] x kept[l]) and B[l] = V_l^+ H V_{l+1}
    coupling levels l and l+1, for l = 1 .. the last level with kept > 0.
    Assembled on the host, reduced on the GPU (Householder, tridiagonal QL)."""
    nl = len(kept) - 1
    for l in range(1, len(kept)):
        if kept[l] == 0:
            nl = l - 1
            break
    sizes = [int(kept[l]) for l in range(1, nl + 1)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    m = int(off[-1])
    T = np.zeros((m, m), dtype=np.complex128)
    for i, l in enumerate(range(1, nl + 1)):
        p = sizes[i]
        T[off[i]:off[i] + p, off[i]:off[i] + p] = A[l][:p, :p]
        if i + 1 < nl:
            q = sizes[i + 1]
            T[off[i]:off[i] + p, off[i + 1]:off[i + 1] + q] = B[l][:p, :q]
            T[off[i + 1]:off[i + 1] + q, off[i]:off[i] + p] = B[l][:p, :q].conj().T
    # a Hermitian matrix read column-major is its conjugate: same eigenvalues
    Hd = torch.from_numpy(T).to(torch.device("cuda", ctx.device)).t()
    E, _ = dense_poles(Hd, None, ctx, vectors=False)
    return E


def _tri_spectra(A, Bv, psiEn, o, ctx):
    """Spectra and shifted Tri tables of Lanczos coefficients (LuaInterface/ssp.cpp:2640-2757)."""
    nsp = A.shape[0]
    gamma, ne = float(o["Gamma"]), int(o["NE"])
    lo, hi = zip(*[_pole_range(A[i], Bv[i], psiEn[i]) for i in range(nsp)])
    eminpole, emaxpole = min(lo), max(hi)
    emin, emax = _default_range(o, eminpole, emaxpole)
    if gamma == 0:
        gamma = 10 * (emax - emin) / ne
    G = cf_spectrum(A, Bv, np.array(psiEn), float(emin), float(emax), ne, gamma, ctx)
    tables = []
    for i in range(nsp):
        a = A[i].copy()
        a[1:] -= psiEn[i]  # ShiftTriDiagonalMatrix shifts a[1..N] (BasicMath_TriDiagonalMatrixVector.cpp:133)
        tables.append(TriResponse(a, Bv[i], 0.0, f"Spectrum {i}"))
    return Spectra(float(emin), float(emax), ne, gamma, G, eminpole, emaxpole), tables


_CS_OPTIONS = {"Tensor": False, "KrylovBasis": False, "NTri": 200, "Epsilon": SQRT_EPS / 100,
               "SingularValue": SQRT_EPS, "Restrictions": None, "Emin": None, "Emax": None, "NE": 1000,
               "Gamma": 0.0, "E0": None, "DenseBorder": 1513, "ForceDenseMethod": False, "Method": None}
_CS_UNSUPPORTED = {n: "bath-excitation, grid and file options of the reference are outside this engine"
                   for n in ("GridPoints", "NTriFull", "NBathExcitations", "NonCrossing", "file")}
QS_KMAX = 16  # vectors in one block of the block Lanczos (qs_block_lanczos)


def _lanczos_many(hop: Hop, vs: list, ntri: int, ctx):
    """Lanczos tables of several start vectors on one operator: all of them as
    slots of one batch, advanced together (one set of launches per step), when
    the binding is assembled; one after another otherwise. Each slot's
    recurrence is that of qs_lanczos (LanczosTriDiagonalizeCC on a fixed basis)."""
    if not vs:
        return []
    if not hop.stats()["assembled"] or len(vs) == 1:
        return [hop.lanczos(v, ntri) for v in vs]
    out = []
    for g in range(0, len(vs), _BATCH_SLOTS):
        part = vs[g:g + _BATCH_SLOTS]
        batch = Batch(len(part), ctx)
        slots = [batch.add(hop, v, ntri) for v in part]
        batch.run(ntri)
        steps, a, b = batch.coeffs(slots, ntri + 1)
        if np.any(steps < ntri):
            raise QSError("batched Lanczos stopped early")
        out += [(a[i, :ntri + 1].copy(), b[i, :ntri].copy()) for i in range(len(part))]
    return out


_BATCH_SLOTS = 64


def CreateSpectra(H: Operator, T, psi, options=None):
    """Spectra of transition operators T on wavefunctions psi (LuaCreateSpectra,
    LuaInterface/ssp.cpp:1906): E0 = <psi|H|psi> unless given, T psi (restricted
    when Restrictions are given), then one of two methods (option Method):
      - dense ("Dense", or "Auto" when the closure of the T psi under H has at
        most DenseBorder determinants, ssp.cpp:2259-2316): the determinant
        matrix of H on that closure, its eigen-decomposition, and the spectra
        from the poles; the tables are lists of poles (KrylovBasis = true adds
        the eigenstates);
      - Lanczos ("Lanczos", or "Auto" above DenseBorder): NTri-step Lanczos per
        (T, psi) — or one block Lanczos over all of them with Tensor = true —
        and the continued fraction; the tables are tridiagonal.
    Spectra on NE+1 points with Lorentzian FWHM Gamma."""
    given: set = set()
    o = _options(options, _CS_OPTIONS, {"epsilon": "Epsilon", "restrictions": "Restrictions"}, _CS_UNSUPPORTED,
                 given=given)
    _flag_untruncated(o, given, "CreateSpectra")
    Ts = list(T) if isinstance(T, (list, tuple)) else [T]
    psis = list(psi) if isinstance(psi, (list, tuple)) else [psi]
    if not Ts or not psis:
        raise QSError("CreateSpectra called with an empty table of operators or wavefunctions")
    nf = H.NF
    # Hermiticity as LuaCreateSpectra tests it (LuaInterface/ssp.cpp:2180-2203), at
    # the default convergence zero (Core/StructDef.cpp:302): an operator that is
    # not Hermitian is an error; one Hermitian up to that zero is used as given
    if hermitian_class(H, 100 * SQRT_EPS) == 0:
        raise QSError("CreateSpectra called with an operator that is not Hermitian")
    if o["Tensor"] and len(Ts) * len(psis) > QS_KMAX:
        raise QSError(f"Tensor CreateSpectra holds {len(Ts) * len(psis)} (T, psi) pairs; the block Lanczos "
                      f"takes at most {QS_KMAX}")
    R = _restrictions(o["Restrictions"], nf)
    ctx = default_context()
    ntri = int(o["NTri"])
    method = str(o["Method"] or "Auto").lower()
    if method not in ("auto", "dense", "lanczos"):
        raise QSError(f'CreateSpectra: unknown Method "{o["Method"]}", possible are "Auto", "Dense" and "Lanczos"')
    if method == "lanczos" and o["ForceDenseMethod"]:
        raise QSError('CreateSpectra: Method "Lanczos" and ForceDenseMethod true contradict each other')
    forced = bool(o["ForceDenseMethod"]) or method == "dense"
    border = 0 if method == "lanczos" else int(o["DenseBorder"])
    E0s = []
    for p in psis:
        E0s.append(float(o["E0"]) if o["E0"] is not None else _expect(H, p).real)
    # T psi for every (T, psi), T outer, psi inner
    tpsi = [operator_psi(t, p, restrictions=[nf, 0, *R] if R else None) for t in Ts for p in psis]
    psiEn = [E0s[i % len(psis)] for i in range(len(tpsi))]
    if forced or border > 0:
        seeds = [tp.host()[0] for tp in tpsi]
        seeds = np.unique(np.concatenate(seeds), axis=0) if any(len(w) for w in seeds) else np.zeros((0, 1))
        if len(seeds) and (forced or len(seeds) <= border):
            seed = Basis.from_words(nf, seeds, ctx)
            cb = Basis.closure(H.device(ctx), seed, [(m, lo, hi) for m, lo, hi in R] if R else None,
                               max_dim=(1 << 62) if forced else border)
            if cb is not None:
                return _dense_spectra(H, cb, tpsi, psiEn, o, ctx)
    # the space H explores from T psi: fix the occupation of every group of modes H conserves
    groups = conserved_groups(H)
    nsp = len(tpsi)
    # an operator that annihilates psi gives an empty T psi: the reference's
    # Lanczos then has b[0] = 0 and the spectrum vanishes (Core/Lanczos.cpp:2189-2191)
    zero = [len(tp.host()[0]) == 0 for tp in tpsi]
    if all(zero):
        A, Bv = np.zeros((nsp, ntri + 1)), np.zeros((nsp, ntri))
        if o["Tensor"]:
            emin = -10.0 if o["Emin"] is None else float(o["Emin"])
            emax = 10.0 if o["Emax"] is None else float(o["Emax"])
            ne = int(o["NE"])
            gamma = float(o["Gamma"]) or 10 * (emax - emin) / ne
            return (Spectra(emin, emax, ne, gamma, np.zeros((nsp * nsp, ne + 1), dtype=np.complex128)),
                    {"A": np.zeros((ntri + 1, nsp, nsp), dtype=np.complex128),
                     "B": np.zeros((ntri, nsp, nsp), dtype=np.complex128), "kept": np.zeros(ntri + 1, np.int32)})
        return _tri_spectra(A, Bv, psiEn, o, ctx)
    cv = set()
    for tp in tpsi:
        w, _ = tp.host()
        cv |= _counts_of_words(nf, groups, w)
    if len(cv) > 1:
        # the T psi fall in different sectors (e.g. T = c+ on modes of two
        # decoupled groups): merge the groups whose occupation differs into one
        # group with a fixed total, a space H leaves invariant that holds every
        # T psi (the extra sectors it may add are not coupled to them)
        vary = [k for k in range(len(groups)) if len({c[k] for c in cv}) > 1]
        totals = {sum(c[k] for k in vary) for c in cv}
        if len(totals) != 1:
            raise QSError("T psi holds different particle numbers on the modes H couples; one sector is supported")
        keep = [k for k in range(len(groups)) if k not in vary]
        c0 = next(iter(cv))
        groups = [groups[k] for k in keep] + [sorted(m for k in vary for m in groups[k])]
        cv = {tuple(c0[k] for k in keep) + (totals.pop(),)}
    rows = _sector_of_counts(nf, groups, cv.pop(), R, ctx)
    for i, tp in enumerate(tpsi):
        w, c = tp.host()
        v = np.zeros(rows.n, dtype=np.complex128)
        if len(w):
            idx = rows.index(w)
            if np.any(idx < 0):
                raise QSError("T psi has determinants outside its sector")
            v[idx] = c
        tpsi[i] = Wavefunction(nf, basis=rows, vec=torch.from_numpy(v).to(torch.device("cuda", ctx.device)))
    hop = Hop(H.device(ctx), rows, mode="auto")
    if not o["Tensor"]:
        A = np.zeros((nsp, ntri + 1))
        Bv = np.zeros((nsp, ntri))
        live = [i for i in range(nsp) if not zero[i]]  # an empty T psi keeps its zero table
        if o["KrylovBasis"]:
            # LanczosTriDiagonalizeKrylov (Core/Lanczos.cpp:3705): fully
            # re-orthogonalized, the Krylov vectors returned as wavefunctions, the
            # empty ones left out (LuaPushTable_WaveFunctionType_RemoveEmpty,
            # LuaInterface/ssp.cpp:2730-2752)
            kb = [[] for _ in range(nsp)]
            for i in live:
                A[i], Bv[i], V = hop.lanczos_krylov(tpsi[i]._vec.contiguous(), ntri, float(o["Epsilon"]))
                kb[i] = [Wavefunction(nf, basis=rows, vec=V[j].clone()) for j in range(ntri)
                         if bool(torch.any(V[j] != 0))]
            sp, tables = _tri_spectra(A, Bv, psiEn, o, ctx)
            return sp, tables, (kb[0] if nsp == 1 else kb)
        coef = _lanczos_many(hop, [tpsi[i]._vec.contiguous() for i in live], ntri, ctx)
        for i, (a, b) in zip(live, coef):
            A[i], Bv[i] = a, b
        return _tri_spectra(A, Bv, psiEn, o, ctx)
    if o["KrylovBasis"]:
        raise QSError("KrylovBasis with Tensor is supported by the dense method only (Method \"Dense\")")
    # tensor: one block Lanczos over all T psi, (nsp x nsp) spectra
    if len(set(psiEn)) != 1:
        raise QSError("Tensor spectra need one E0 for all wavefunctions (pass E0)")
    V0 = torch.stack([tp._vec for tp in tpsi]).contiguous()
    A, B, kept = hop.block_lanczos(V0, ntri, float(o["SingularValue"]))
    ne = int(o["NE"])
    # the pole range is the extreme eigenvalues of the block tridiagonal matrix,
    # reported whether or not Emin/Emax are given, and the default energy range
    # follows from it (LuaInterface/ssp.cpp:2824-2846, tensor Lanczos branch)
    ev = _block_tridiag_eigenvalues(A, B, kept, ctx)
    eminpole, emaxpole = ev[0] - psiEn[0], ev[-1] - psiEn[0]
    emin, emax = _default_range(o, eminpole, emaxpole)
    gamma = float(o["Gamma"]) or 10 * (emax - emin) / ne
    G = block_cf_spectrum(A, B, kept, psiEn[0], emin, emax, ne, gamma, ctx)
    return (Spectra(emin, emax, ne, gamma, G.reshape(ne + 1, -1).T, eminpole, emaxpole),
            {"A": A, "B": B, "kept": kept})
