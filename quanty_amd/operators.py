"""Operator construction: ladder-term lists for the atomic and impurity
Hamiltonians used by the spectra path.

A term list is [(codes, value), ...]; codes are uint16 ladder codes with o[0]
applied first (bit 15 = create), value a complex number. These builders cover
the named operators the workloads need (NewOperator "U", "ldots", "Number",
"Cr"/"An", one-particle matrices, Oh crystal field; LuaInterface/Operator.cpp,
AtomicOperators/), written from the Slater-Condon rules:
  <m1 m2|1/r12|m3 m4> = sum_k R^k c^k(l1 m1, l3 m3) c^k(l4 m4, l2 m2)
  c^k(l m, l' m') = (-1)^m sqrt((2l+1)(2l'+1)) (l k l'; 0 0 0)(l k l'; -m m-m' m')
Shell modes are given as IndexUp/IndexDn lists, entry j holding m = -l + j.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

CR = 0x8000


def cr(i: int) -> int:
    return CR | i


def an(i: int) -> int:
    return i


def written(ops, value):
    """Term from a written product, e.g. [cr(a), an(b)] = c+_a c_b (rightmost acts first)."""
    return (tuple(reversed(ops)), complex(value))


# ---------------------------------------------------------------------------
# angular algebra
# ---------------------------------------------------------------------------
@lru_cache(maxsize=None)
def wigner3j(j1, j2, j3, m1, m2, m3) -> float:
    """Wigner 3j symbol for integer arguments (Racah's formula)."""
    if m1 + m2 + m3 != 0:
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    f = math.factorial
    tri = f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3) / f(j1 + j2 + j3 + 1)
    pre = math.sqrt(tri * f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j3 + m3) * f(j3 - m3))
    s = 0.0
    for t in range(0, j1 + j2 + j3 + 1):
        a = [t, j3 - j2 + t + m1, j3 - j1 + t - m2, j1 + j2 - j3 - t, j1 - t - m1, j2 - t + m2]
        if min(a) < 0:
            continue
        s += (-1) ** t / (f(a[0]) * f(a[1]) * f(a[2]) * f(a[3]) * f(a[4]) * f(a[5]))
    return (-1) ** (j1 - j2 - m3) * pre * s


@lru_cache(maxsize=None)
def ck(k, l1, m1, l2, m2) -> float:
    """Gaunt coefficient c^k(l1 m1, l2 m2) = <l1 m1|C^k_{m1-m2}|l2 m2>."""
    return ((-1) ** m1 * math.sqrt((2 * l1 + 1) * (2 * l2 + 1)) * wigner3j(l1, k, l2, 0, 0, 0)
            * wigner3j(l1, k, l2, -m1, m1 - m2, m2))


def _shell(index_up, index_dn):
    l = (len(index_up) - 1) // 2
    assert len(index_up) == len(index_dn) == 2 * l + 1
    orbs = []  # (mode, m, spin)
    for j in range(2 * l + 1):
        orbs.append((index_dn[j], j - l, -1))
        orbs.append((index_up[j], j - l, +1))
    return l, orbs


# ---------------------------------------------------------------------------
# two-body
# ---------------------------------------------------------------------------
def coulomb(index_up, index_dn, F):
    """Intra-shell Coulomb with Slater integrals F = [F0, F2, F4, ...] (NewOperator "U")."""
    l, orbs = _shell(index_up, index_dn)
    ks = [2 * i for i in range(len(F))]
    out = {}
    for (i1, m1, s1) in orbs:
        for (i2, m2, s2) in orbs:
            if i1 == i2:
                continue
            for (i3, m3, s3) in orbs:
                if s3 != s1:
                    continue
                for (i4, m4, s4) in orbs:
                    if s4 != s2 or i3 == i4 or m1 + m2 != m3 + m4:
                        continue
                    v = sum(Fk * ck(k, l, m1, l, m3) * ck(k, l, m4, l, m2) for k, Fk in zip(ks, F))
                    if abs(v) < 1e-15:
                        continue
                    key = (an(i3), an(i4), cr(i2), cr(i1))  # c+_1 c+_2 c_4 c_3
                    out[key] = out.get(key, 0.0) + 0.5 * v
    return [(k, complex(v)) for k, v in out.items()]


def coulomb_two_shells(up_a, dn_a, up_b, dn_b, F, G):
    """Inter-shell Coulomb between shells A and B: direct F^k (k = 0, 2, ...) and
    exchange G^k (k = |la-lb|, |la-lb|+2, ...) (NewOperator "U" with two shells)."""
    la, oa = _shell(up_a, dn_a)
    lb, ob = _shell(up_b, dn_b)
    orbs = [(i, m, s, 0, la) for (i, m, s) in oa] + [(i, m, s, 1, lb) for (i, m, s) in ob]
    kF = [2 * i for i in range(len(F))]
    kG = [abs(la - lb) + 2 * i for i in range(len(G))]
    out = {}
    for (i1, m1, s1, h1, l1) in orbs:
        for (i2, m2, s2, h2, l2) in orbs:
            if i1 == i2 or h1 == h2:
                continue
            for (i3, m3, s3, h3, l3) in orbs:
                if s3 != s1:
                    continue
                for (i4, m4, s4, h4, l4) in orbs:
                    if s4 != s2 or i3 == i4 or m1 + m2 != m3 + m4:
                        continue
                    if h3 == h1 and h4 == h2:
                        ks, R = kF, F
                    elif h3 == h2 and h4 == h1:
                        ks, R = kG, G
                    else:
                        continue
                    v = sum(Rk * ck(k, l1, m1, l3, m3) * ck(k, l4, m4, l2, m2) for k, Rk in zip(ks, R))
                    if abs(v) < 1e-15:
                        continue
                    key = (an(i3), an(i4), cr(i2), cr(i1))
                    out[key] = out.get(key, 0.0) + 0.5 * v
    return [(k, complex(v)) for k, v in out.items()]


# ---------------------------------------------------------------------------
# one-body
# ---------------------------------------------------------------------------
@lru_cache(maxsize=1024)
def _one_body_codes(rows: tuple, cols: tuple) -> tuple:
    """Ladder codes (o[0] first) of c+_rows[i] c_cols[j], row-major over (i, j)."""
    return tuple((an(b), cr(a)) for a in rows for b in cols)


def one_body(rows, cols, M):
    """sum_ij M[i][j] c+_rows[i] c_cols[j] (zero entries left out)."""
    M = np.asarray(M, dtype=np.complex128).reshape(len(rows), len(cols))
    codes = _one_body_codes(tuple(int(a) for a in rows), tuple(int(b) for b in cols))
    flat = M.ravel()
    nz = np.flatnonzero(flat)
    return [(codes[k], complex(flat[k])) for k in nz.tolist()]


def number(indices, values=None):
    values = [1.0] * len(indices) if values is None else values
    return [written([cr(i), an(i)], v) for i, v in zip(indices, values) if v != 0]


def _lmat(l):
    """Lz, L+ in the m = -l..l basis."""
    d = 2 * l + 1
    Lz = np.diag([float(m) for m in range(-l, l + 1)])
    Lp = np.zeros((d, d))
    for j in range(d - 1):
        m = j - l
        Lp[j + 1, j] = math.sqrt(l * (l + 1) - m * (m + 1))
    return Lz, Lp


def spin_orbital_matrix(l, which):
    """(2(2l+1))^2 one-particle matrix in the (m, spin) basis ordered dn_m, up_m per m."""
    d = 2 * l + 1
    Lz, Lp = _lmat(l)
    Lm = Lp.T
    I2 = np.eye(2)
    Id = np.eye(d)
    sz = np.diag([-0.5, 0.5])            # dn, up
    sp = np.array([[0.0, 0.0], [1.0, 0.0]])  # s+ |dn> = |up>
    sm = sp.T
    kron = np.kron
    if which == "Lz":
        return kron(Lz, I2)
    if which == "Lx":
        return kron(0.5 * (Lp + Lm), I2)
    if which == "Ly":
        return kron(-0.5j * (Lp - Lm), I2)
    if which == "Sz":
        return kron(Id, sz)
    if which == "Sx":
        return kron(Id, 0.5 * (sp + sm))
    if which == "Sy":
        return kron(Id, -0.5j * (sp - sm))
    if which == "ldots":
        return kron(Lz, sz) + 0.5 * (kron(Lp, sm) + kron(Lm, sp))
    if which == "Splus":
        return kron(Id, sp)
    if which == "Smin":
        return kron(Id, sm)
    if which == "Lplus":
        return kron(Lp, I2)
    if which == "Lmin":
        return kron(Lm, I2)
    if which in ("Jx", "Jy", "Jz", "Jplus", "Jmin"):
        s_, l_ = {"Jx": ("Sx", "Lx"), "Jy": ("Sy", "Ly"), "Jz": ("Sz", "Lz"), "Jplus": ("Splus", "Lplus"),
                  "Jmin": ("Smin", "Lmin")}[which]
        return spin_orbital_matrix(l, s_) + spin_orbital_matrix(l, l_)
    raise ValueError(which)


def shell_operator(index_up, index_dn, which, scale=1.0):
    l, orbs = _shell(index_up, index_dn)
    modes = [o[0] for o in orbs]
    return one_body(modes, modes, scale * spin_orbital_matrix(l, which))


def oh_d_projectors():
    """eg and t2g projectors of a d shell in the m = -2..2 basis (cubic harmonics:
    eg = {Y20, (Y22+Y2-2)/sqrt2}, t2g = {Y21, Y2-1, (Y22-Y2-2)/sqrt2})."""
    s = 1 / math.sqrt(2)
    eg = np.zeros((5, 5))
    eg[2, 2] = 1.0
    v = np.zeros(5)
    v[0] = v[4] = s
    eg += np.outer(v, v)
    t2g = np.eye(5) - eg
    return eg, t2g


def d_shell_matrix_to_spin(M5):
    """5x5 orbital matrix -> 10x10 spin-diagonal (dn, up per m)."""
    return np.kron(np.asarray(M5), np.eye(2))


def shell_modes(index_up, index_dn):
    _, orbs = _shell(index_up, index_dn)
    return [o[0] for o in orbs]


def scale(terms, f):
    return [(c, complex(v) * f) for c, v in terms]


def add(*lists):
    out = {}
    for lst in lists:
        for c, v in lst:
            out[c] = out.get(c, 0.0) + complex(v)
    return [(c, v) for c, v in out.items() if v != 0]


@lru_cache(maxsize=1 << 16)
def _ordered(codes: tuple) -> tuple:
    """A ladder string as a sum of strings in canonical order, codes strictly
    ascending in application order (annihilators before creators, each by
    ascending mode; written: creators left of annihilators). Adjacent factors
    anticommute, {c_i, c+_i} = 1 adds the string without the pair, and a
    repeated code makes the string vanish. Returns ((codes, coefficient), ...)."""
    for k in range(len(codes) - 1):
        a, b = codes[k], codes[k + 1]
        if a == b:
            return ()
        if a > b:
            out: dict = {}
            for c, v in _ordered(codes[:k] + (b, a) + codes[k + 2:]):
                out[c] = out.get(c, 0) - v
            if a ^ b == CR:
                for c, v in _ordered(codes[:k] + codes[k + 2:]):
                    out[c] = out.get(c, 0) + v
            return tuple((c, v) for c, v in out.items() if v != 0)
    return ((codes, 1),)


def normal_order(terms, eps: float = 0.0) -> dict:
    """Terms rewritten as a sum of canonically ordered strings with their
    coefficients added, every coefficient with |v| <= eps dropped — the
    normal-ordered form the reference keeps operators in
    (OperatorAddLaderNormalOrder, Core/OperatorArithmetic.cpp:1212) and the
    chop of OperatorRemoveSmallValues (:1309-1420; |v| > eps is kept). The
    canonical order here is this module's own; sums that cancel cancel in any
    canonical order."""
    out: dict = {}
    items = terms.items() if isinstance(terms, dict) else terms
    for c, v in items:
        for cc, f in _ordered(tuple(int(x) for x in c)):
            out[cc] = out.get(cc, 0) + f * complex(v)
    return {c: v for c, v in out.items() if abs(v) > eps}


def dagger(terms):
    """Conjugate transpose: reverse each string, swap c+/c, conjugate."""
    return [(tuple(x ^ CR for x in reversed(c)), complex(v).conjugate()) for c, v in terms]


def dipole_p_to_d(up_d, dn_d, up_p, dn_p, akm):
    """Dipole-type transition sum_{k,q} A_kq <d m|C^k_q|p m'> c+_{d m s} c_{p m' s} ("CF" between shells)."""
    ld, od = _shell(up_d, dn_d)
    lp, op = _shell(up_p, dn_p)
    out = {}
    for (k, q, A) in akm:
        for (i1, m1, s1) in od:
            for (i2, m2, s2) in op:
                if s1 != s2 or m1 != q + m2:
                    continue
                v = A * ck(k, ld, m1, lp, m2)
                if v == 0:
                    continue
                key = (an(i2), cr(i1))
                out[key] = out.get(key, 0.0) + v
    return [(c, complex(v)) for c, v in out.items()]
