"""Seeded impurity problems for the spectra path.

* ``ni_xas``      — Ni2+ O_h charge-transfer L2,3 XAS: 26 modes (2p 0-5, 3d 6-15,
  ligand 16-25, dn/up interleaved), initial and final Hamiltonians following the
  ligand-field tutorial (Debug examples/Tutorials/40_NiO_Ligand_Field/41_XAS.Quanty)
  with the parameters of the workload; ground sector 2p6 (d+L)18 (190 dets),
  final sector 2p5 (d+L)19 (120 dets).
* ``anderson``    — Anderson impurity: 3d shell plus bath shells of 10 modes each,
  dd Coulomb, SOC and O_h field on d, random bath energies U(-5,5), complex
  d-bath hybridisations |V|~U(0.5,2), complex Gaussian intra-bath hopping
  sigma=0.3, occupation restrictions per shell group.
All randomness comes from SplitMix64 streams so a seed fixes the problem.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

from . import operators as O

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 stream (Steele, Lea, Flood 2014); uniform doubles in [0,1)."""

    def __init__(self, seed: int):
        self.s = seed & MASK64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0) -> float:
        return a + (b - a) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def gauss(self, sigma=1.0) -> float:
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return sigma * math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def cgauss(self, sigma=1.0) -> complex:
        """Complex Gaussian with E|z|^2 = sigma^2."""
        return complex(self.gauss(sigma / math.sqrt(2)), self.gauss(sigma / math.sqrt(2)))

    def uniforms(self, n: int, a=0.0, b=1.0) -> np.ndarray:
        """The next n values of `uniform(a, b)` in one vectorized draw."""
        k = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.s) + k * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.s = (self.s + n * 0x9E3779B97F4A7C15) & MASK64
        return a + (b - a) * ((z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0))

    def cgauss_array(self, n: int, sigma=1.0) -> np.ndarray:
        """The next n values of `cgauss(sigma)` (4 uniforms each, in the same order)."""
        u = self.uniforms(4 * n).reshape(n, 4)
        s = sigma / math.sqrt(2)
        re = s * np.sqrt(-2.0 * np.log(np.maximum(u[:, 0], 1e-300))) * np.cos(2.0 * math.pi * u[:, 1])
        im = s * np.sqrt(-2.0 * np.log(np.maximum(u[:, 2], 1e-300))) * np.cos(2.0 * math.pi * u[:, 3])
        return re + 1j * im


def shell(first: int, l: int = 2):
    """(IndexUp, IndexDn) of a shell starting at mode `first`, dn/up interleaved."""
    n = 2 * l + 1
    dn = [first + 2 * j for j in range(n)]
    up = [first + 2 * j + 1 for j in range(n)]
    return up, dn


def mask(nf: int, modes) -> str:
    s = ["0"] * nf
    for i in modes:
        s[i] = "1"
    return "".join(s)


# ---------------------------------------------------------------------------
# Ni2+ L2,3 XAS (charge transfer)
# ---------------------------------------------------------------------------
NI_PARAMS = dict(F2dd=9.79, F4dd=6.08, F2pd=6.18, G1pd=4.63, G3pd=2.63, zeta_3d=0.083, zeta_2p=11.51,
                 tenDq=1.1, tenDqL=1.44, Delta=4.0, Udd=7.3, Upd=8.5, Veg=2.06, Vt2g=1.21, nd=8)


def ni_xas(params=None):
    p = dict(NI_PARAMS)
    p.update(params or {})
    NF = 26
    up_p, dn_p = shell(0, 1)
    up_d, dn_d = shell(6, 2)
    up_L, dn_L = shell(16, 2)
    nd = p["nd"]
    Udd, Upd, Delta = p["Udd"], p["Upd"], p["Delta"]
    ed = (10 * Delta - nd * (19 + nd) * Udd / 2) / (10 + nd)
    eL = nd * ((1 + nd) * Udd / 2 - Delta) / (10 + nd)
    epf = (10 * Delta + (1 + nd) * (nd * Udd / 2 - (10 + nd) * Upd)) / (16 + nd)
    edf = (10 * Delta - nd * (31 + nd) * Udd / 2 - 90 * Upd) / (16 + nd)
    eLf = ((1 + nd) * (nd * Udd / 2 + 6 * Upd) - (6 + nd) * Delta) / (16 + nd)
    F0dd = Udd + (p["F2dd"] + p["F4dd"]) * 2 / 63
    F0pd = Upd + p["G1pd"] / 15 + 3 * p["G3pd"] / 70
    eg, t2g = O.oh_d_projectors()
    dmodes = O.shell_modes(up_d, dn_d)
    Lmodes = O.shell_modes(up_L, dn_L)
    pmodes = O.shell_modes(up_p, dn_p)
    cf = O.d_shell_matrix_to_spin(0.6 * eg - 0.4 * t2g)
    hyb = O.d_shell_matrix_to_spin(p["Veg"] * eg + p["Vt2g"] * t2g)
    common = O.add(
        O.coulomb(up_d, dn_d, [F0dd, p["F2dd"], p["F4dd"]]),
        O.shell_operator(up_d, dn_d, "ldots", p["zeta_3d"]),
        O.one_body(dmodes, dmodes, p["tenDq"] * cf),
        O.one_body(Lmodes, Lmodes, p["tenDqL"] * cf),
        O.one_body(dmodes, Lmodes, hyb),
        O.one_body(Lmodes, dmodes, hyb.conj().T),
    )
    Hi = O.add(common, O.number(dmodes, [ed] * 10), O.number(Lmodes, [eL] * 10))
    Hf = O.add(common, O.number(dmodes, [edf] * 10), O.number(Lmodes, [eLf] * 10), O.number(pmodes, [epf] * 6),
               O.shell_operator(up_p, dn_p, "ldots", p["zeta_2p"]),
               O.coulomb_two_shells(up_p, dn_p, up_d, dn_d, [F0pd, p["F2pd"]], [p["G1pd"], p["G3pd"]]))
    t = math.sqrt(0.5)
    Tx = O.dipole_p_to_d(up_d, dn_d, up_p, dn_p, [(1, -1, t), (1, 1, -t)])
    Ty = O.dipole_p_to_d(up_d, dn_d, up_p, dn_p, [(1, -1, 1j * t), (1, 1, 1j * t)])
    Tz = O.dipole_p_to_d(up_d, dn_d, up_p, dn_p, [(1, 0, 1.0)])
    gs = [(mask(NF, pmodes), 6, 6), (mask(NF, dmodes + Lmodes), 18, 18)]
    fs = [(mask(NF, pmodes), 5, 5), (mask(NF, dmodes + Lmodes), 19, 19)]
    return dict(NF=NF, Hi=Hi, Hf=Hf, T=[Tx, Ty, Tz], gs_restrictions=gs, fs_restrictions=fs, params=p)


# ---------------------------------------------------------------------------
# Anderson impurity: d shell + bath shells
# ---------------------------------------------------------------------------
def anderson_restrictions(nbath: int = 4, nelec: int = 28, full_baths: int = 2, d_range=(6, 10),
                          max_holes: int = 2, max_elec: int = 2):
    """Restrictions of the Anderson sector: total N, d occupation window, at most
    `max_holes` holes in the nominally full baths, at most `max_elec` electrons in
    the nominally empty ones."""
    NF = 10 * (1 + nbath)
    dmodes = list(range(10))
    full = list(range(10, 10 + 10 * full_baths))
    empty = list(range(10 + 10 * full_baths, NF))
    restr = [(mask(NF, range(NF)), nelec, nelec), (mask(NF, dmodes), d_range[0], d_range[1])]
    if full:
        restr.append((mask(NF, full), len(full) - max_holes, len(full)))
    if empty:
        restr.append((mask(NF, empty), 0, max_elec))
    return restr


@lru_cache(maxsize=8)
def _impurity_d_terms(coulomb, zeta, tenDq):
    up_d, dn_d = shell(0, 2)
    dmodes = O.shell_modes(up_d, dn_d)
    Udd, F2, F4 = coulomb
    F0 = Udd + (F2 + F4) * 2 / 63
    eg, t2g = O.oh_d_projectors()
    return tuple(O.coulomb(up_d, dn_d, [F0, F2, F4]) + O.shell_operator(up_d, dn_d, "ldots", zeta)
                 + O.one_body(dmodes, dmodes, tenDq * O.d_shell_matrix_to_spin(0.6 * eg - 0.4 * t2g)))


_IU = np.triu_indices(10, 1)


def _bath_couplings(rng: SplitMix64):
    """One bath shell's random parameters, drawn in this order: 10 energies
    U(-5, 5); for each orbital a hybridization r e^{i phi}, r U(0.5, 2),
    phi U(0, 2 pi); the intra-bath hoppings T_ij (i < j) complex Gaussian of
    width 0.3, row by row, with T_ji = conj(T_ij)."""
    eps = rng.uniforms(10, -5.0, 5.0)
    u = rng.uniforms(20)
    r = 0.5 + (2.0 - 0.5) * u[0::2]
    ph = 0.0 + (2 * math.pi - 0.0) * u[1::2]
    V = np.diag(r * (np.cos(ph) + 1j * np.sin(ph)))
    T = np.zeros((10, 10), dtype=complex)
    z = rng.cgauss_array(45, 0.3)
    T[_IU] = z
    T[(_IU[1], _IU[0])] = z.conj()
    return eps, V, T


def anderson(seed: int = 1, nbath: int = 4, nelec: int = 28, full_baths: int = 2, d_range=(6, 10),
             max_holes: int = 2, max_elec: int = 2, coulomb=(7.3, 9.79, 6.08), zeta=0.083, tenDq=1.1):
    """Anderson impurity with `nbath` bath shells; the first `full_baths` are
    nominally full (at most `max_holes` holes in total), the rest nominally empty
    (at most `max_elec` electrons in total)."""
    rng = SplitMix64(seed)
    NF = 10 * (1 + nbath)
    up_d, dn_d = shell(0, 2)
    dmodes = O.shell_modes(up_d, dn_d)
    terms = [list(_impurity_d_terms(tuple(coulomb), zeta, tenDq))]
    baths = []
    for b in range(nbath):
        up_b, dn_b = shell(10 + 10 * b, 2)
        bm = O.shell_modes(up_b, dn_b)
        baths.append(bm)
        eps, V, T = _bath_couplings(rng)
        terms.append(O.number(bm, list(eps)))
        # d <-> bath, orbital diagonal (same m and spin)
        terms.append(O.one_body(dmodes, bm, V))
        terms.append(O.one_body(bm, dmodes, V.conj().T))
        # intra-bath hopping, Hermitian, off-diagonal
        terms.append(O.one_body(bm, bm, T))
    H = O.add(*terms)
    restr = anderson_restrictions(nbath, nelec, full_baths, d_range, max_holes, max_elec)
    return dict(NF=NF, H=H, restrictions=restr, dmodes=dmodes, baths=baths, seed=seed)


def random_one_body_transition(rng: SplitMix64, dmodes, bath_modes, sigma=1.0):
    """Random one-body d<->bath transition: sum t_ij c+_i c_j over d->bath and
    bath->d pairs, every t_ij an independent complex Gaussian of width sigma."""
    M1 = rng.cgauss_array(len(dmodes) * len(bath_modes), sigma).reshape(len(dmodes), len(bath_modes))
    M2 = rng.cgauss_array(len(bath_modes) * len(dmodes), sigma).reshape(len(bath_modes), len(dmodes))
    return O.add(O.one_body(dmodes, bath_modes, M1), O.one_body(bath_modes, dmodes, M2))


# ---------------------------------------------------------------------------
# 76 modes (two-word determinants): 2p core + d shell + 6 bath shells
# ---------------------------------------------------------------------------
def large76(seed: int = 1, params=None, nd: int = 8):
    """L2,3 edge of a d impurity with six bath shells: 2p core (modes 0-5), 3d
    (6-15), B1-B3 nominally full (16-45), B4-B6 nominally empty (46-75), N = 44.
    Sectors: 2p6 (ground state, 10,347,855 dets) and 2p5 (15,330,840 dets), both
    with d occupation 6-10, at most 2 holes in B1-B3 and at most 2 electrons in
    B4-B6. H_i: dd Coulomb, 3d spin-orbit and 10Dq of the Ni2+ parameters, bath
    energies U(-5, 5), d<->bath hybridizations r e^{i phi} (r U(0.5, 2)) diagonal
    in the orbital, complex Gaussian intra-bath hoppings (sigma 0.3). H_f adds
    the 2p-3d Coulomb (F0, F2; G1, G3), 2p spin-orbit and a core energy. T is
    the z-polarized 2p -> 3d dipole."""
    p = dict(NI_PARAMS)
    p.update(params or {})
    rng = SplitMix64(seed)
    NF = 76
    up_p, dn_p = shell(0, 1)
    up_d, dn_d = shell(6, 2)
    pmodes = O.shell_modes(up_p, dn_p)
    dmodes = O.shell_modes(up_d, dn_d)
    F0dd = p["Udd"] + (p["F2dd"] + p["F4dd"]) * 2 / 63
    F0pd = p["Upd"] + p["G1pd"] / 15 + 3 * p["G3pd"] / 70
    eg, t2g = O.oh_d_projectors()
    terms = [O.coulomb(up_d, dn_d, [F0dd, p["F2dd"], p["F4dd"]]),
             O.shell_operator(up_d, dn_d, "ldots", p["zeta_3d"]),
             O.one_body(dmodes, dmodes, p["tenDq"] * O.d_shell_matrix_to_spin(0.6 * eg - 0.4 * t2g)),
             O.number(dmodes, [-nd * p["Udd"] / 2] * 10)]
    baths = []
    for b in range(6):
        up_b, dn_b = shell(16 + 10 * b, 2)
        bm = O.shell_modes(up_b, dn_b)
        baths.append(bm)
        eps, V, T = _bath_couplings(rng)
        terms.append(O.number(bm, list(eps)))
        terms.append(O.one_body(dmodes, bm, V))
        terms.append(O.one_body(bm, dmodes, V.conj().T))
        terms.append(O.one_body(bm, bm, T))
    Hi = O.add(*terms)
    Hf = O.add(Hi, O.shell_operator(up_p, dn_p, "ldots", p["zeta_2p"]), O.number(pmodes, [-p["Upd"] * 2] * 6),
               O.coulomb_two_shells(up_p, dn_p, up_d, dn_d, [F0pd, p["F2pd"]], [p["G1pd"], p["G3pd"]]))
    Tz = O.dipole_p_to_d(up_d, dn_d, up_p, dn_p, [(1, 0, 1.0)])
    full = [m for bm in baths[:3] for m in bm]
    empty = [m for bm in baths[3:] for m in bm]
    rest = [(mask(NF, dmodes), 6, 10), (mask(NF, full), 28, 30), (mask(NF, empty), 0, 2)]
    gs = [(mask(NF, range(NF)), 44, 44), (mask(NF, pmodes), 6, 6)] + rest
    fs = [(mask(NF, range(NF)), 44, 44), (mask(NF, pmodes), 5, 5)] + rest
    return dict(NF=NF, Hi=Hi, Hf=Hf, T=Tz, gs_restrictions=gs, fs_restrictions=fs, dmodes=dmodes, baths=baths,
                pmodes=pmodes, seed=seed)
