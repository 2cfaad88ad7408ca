"""Small models shared by the tests (mirroring the reference's test scripts)."""
import json
import math
import os

import numpy as np

from quanty_amd import operators as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    return json.load(open(os.path.join(GOLDEN, name)))


def chain(nbath=5, kind=0, U=3.0, t=1.0):
    """Impurity (modes 0,1) + chain of `nbath` sites (TestScripts/CreateSpectra.Quanty:269-300).
    kind 0: U n0 n1 + real hopping; 1: U n0 n1 + imaginary hopping; 2: imaginary hopping only."""
    nf = 2 + 2 * nbath
    terms = []
    if kind in (0, 1):
        terms.append(O.written([O.cr(0), O.an(0), O.cr(1), O.an(1)], U))
    h = t if kind == 0 else 1j * t
    for i in range(1, nbath + 1):
        terms.append(O.written([O.cr(2 * i), O.an(2 * (i - 1))], h))
        terms.append(O.written([O.cr(2 * i + 1), O.an(2 * (i - 1) + 1)], h))
        terms.append(O.written([O.cr(2 * (i - 1)), O.an(2 * i)], np.conj(h)))
        terms.append(O.written([O.cr(2 * (i - 1) + 1), O.an(2 * i + 1)], np.conj(h)))
    return nf, terms


def chain_states(nf, which):
    """psi1 = (|10..>+|01..>)/sqrt2, psi2 = (|10..> - 2i|01..>)/sqrt5 as (strings, coefficients)."""
    s1 = "10" + "00" * ((nf - 2) // 2)
    s2 = "01" + "00" * ((nf - 2) // 2)
    if which == 1:
        return [s1, s2], np.array([1, 1]) / math.sqrt(2)
    return [s1, s2], np.array([1, -2j]) / math.sqrt(5)


def f_shell(complex_case=False):
    up = [1, 3, 5, 7, 9, 11, 13]
    dn = [0, 2, 4, 6, 8, 10, 12]
    terms = O.add(O.coulomb(up, dn, [0, 10, 6.25, 4]), O.shell_operator(up, dn, "ldots", 0.1))
    if complex_case:
        terms = O.add(terms, O.shell_operator(up, dn, "Ly", 1e-6), O.shell_operator(up, dn, "Sy", 2e-6))
    return 14, terms


def str_to_bytes(s, nf):
    b = np.zeros((nf + 7) // 8, dtype=np.uint8)
    bits = [c for c in s if c in "01"]
    for i, c in enumerate(bits):
        if c == "1":
            b[i >> 3] |= 128 >> (i & 7)
    return b


def jw_matrices(nf):
    """Dense annihilators c_i on the 2^nf Fock space by Jordan-Wigner Kronecker
    products (mode 0 leftmost), basis index = integer with mode 0 as the most
    significant bit — the same order as the numeric determinant order."""
    a = np.array([[0.0, 1.0], [0.0, 0.0]])  # |0><1| on one mode (state index 1 = occupied)
    Z = np.diag([1.0, -1.0])
    I = np.eye(2)
    out = []
    for i in range(nf):
        m = np.array([[1.0]])
        for j in range(nf):
            m = np.kron(m, Z if j < i else (a if j == i else I))
        out.append(m)
    return out


def dense_from_terms(nf, terms, cs=None):
    cs = jw_matrices(nf) if cs is None else cs
    dim = 2 ** nf
    H = np.zeros((dim, dim), dtype=complex)
    for codes, v in terms:
        M = np.eye(dim)
        for c in codes:  # o[0] acts first: M = op_k ... op_0
            op = cs[c & 0x7FFF].T if c & 0x8000 else cs[c & 0x7FFF]
            M = op @ M
        H += v * M
    return H


def hubbard_ring(nsites=8, t1=1.0, t2=0.1, U=3.5):
    """1D Hubbard ring (TestScripts/Eigensystem.Quanty:1121-1136): modes 2i = site i
    down, 2i+1 = site i up; +t1 between nearest and +t2 between next-nearest
    neighbours (periodic, spin-diagonal), U n_dn n_up on every site."""
    nf = 2 * nsites
    terms = []
    for i in range(nsites):
        for d, t in ((1, t1), (2, t2)):
            j = (i + d) % nsites
            for s in (0, 1):
                a, b = 2 * i + s, 2 * j + s
                terms.append(O.written([O.cr(b), O.an(a)], t))
                terms.append(O.written([O.cr(a), O.an(b)], t))
        terms.append(O.written([O.cr(2 * i), O.an(2 * i), O.cr(2 * i + 1), O.an(2 * i + 1)], U))
    return nf, terms
