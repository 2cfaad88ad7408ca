"""Crystal-field potentials expanded on renormalized spherical harmonics:
V = sum_{k,m} A_km C^k_m, restricted to an l shell, such that V has the given
energies on the irreducible representations of a point group (the input of
NewOperator("CF", ...); PotentialExpandedOnClm, LuaInterface/ssp.cpp:908,
tables of Core/BasicMath/BasicMath_StandardFunctions.cpp:1182-4700).

Built here from the group theory instead of per-group formulas: each energy
parameter of a group is an operator on the l shell, either the projector on
the functions of one irreducible representation or, for a parameter that
mixes two equivalent representations, |a><b| + |b><a| summed over partners.
The potential sum_p E_p O_p is then expanded on the matrices
<l m1|C^k_m|l m2> (k even up to 2l), which span the shell's operators of even
parity, by solving the linear system exactly. The energies are given in the
order and orientation of the reference's tables: the highest-order axis along
z, x along a two-fold axis, functions named by real harmonics (z^2, x^2-y^2,
xy, xz, yz)."""
from __future__ import annotations

import math

import numpy as np

from .operators import ck

S2 = math.sqrt(2.0)


def _ket(l, coeffs):
    """A vector of the l shell in the m basis (m = -l..l) from {m: coefficient}."""
    v = np.zeros(2 * l + 1, dtype=complex)
    for m, c in coeffs.items():
        v[m + l] += c
    return v


def real_harmonics(l: int) -> dict:
    """Real (tesseral) harmonics of the l shell as vectors in the complex m basis
    with the Condon-Shortley phase: c_m = ((-1)^m |m> + |-m>)/sqrt2 and
    s_m = i(|-m> - (-1)^m |m>)/sqrt2 for m > 0, and |0>. Named for l <= 2."""
    out = {"m0": _ket(l, {0: 1.0})}
    for m in range(1, l + 1):
        out[f"c{m}"] = _ket(l, {m: (-1) ** m / S2, -m: 1 / S2})
        out[f"s{m}"] = _ket(l, {m: -1j * (-1) ** m / S2, -m: 1j / S2})
    if l == 1:
        out.update(z=out["m0"], x=out["c1"], y=out["s1"])
    if l == 2:
        out.update(z2=out["m0"], xz=out["c1"], yz=out["s1"], x2y2=out["c2"], xy=out["s2"])
    return out


def _projector(vecs):
    P = 0
    for v in vecs:
        v = np.asarray(v, dtype=complex)
        P = P + np.outer(v, v.conj()) / np.vdot(v, v).real
    return P


def _mix(pairs):
    M = 0
    for a, b in pairs:
        M = M + np.outer(a, b.conj()) + np.outer(b, a.conj())
    return M


def _cubic_d():
    R = real_harmonics(2)
    # eg = {z^2, x^2-y^2}, t2g = {xy, xz, yz}
    return [_projector([R["z2"], R["x2y2"]]), _projector([R["xy"], R["xz"], R["yz"]])]


def _tetragonal(l):
    R = real_harmonics(l)
    if l == 1:  # a (z), e (x, y)
        return [_projector([R["z"]]), _projector([R["x"], R["y"]])]
    # a1 (z^2), b1 (x^2-y^2), b2 (xy), e (xz, yz)
    return [_projector([R["z2"]]), _projector([R["x2y2"]]), _projector([R["xy"]]),
            _projector([R["xz"], R["yz"]])]


def _d2h():
    R = real_harmonics(2)
    # ag (x^2-y^2), ag (z^2), their mixing, b1g (xy), b2g (xz), b3g (yz)
    return [_projector([R["x2y2"]]), _projector([R["z2"]]), _mix([(R["x2y2"], R["z2"])]),
            _projector([R["xy"]]), _projector([R["xz"]]), _projector([R["yz"]])]


def _group_ops(group: str, l: int):
    """The operators of the energy parameters of `group` on the l shell, in the
    reference's order, or None when not tabulated here."""
    if group in ("C1", "Ci") or l == 0:
        return [np.eye(2 * l + 1, dtype=complex)]
    if group in ("Oh", "O", "Td") and l == 1:
        return [np.eye(3, dtype=complex)]
    if group in ("Oh", "O", "Td") and l == 2:
        return _cubic_d()
    if group in ("D4h", "C4v", "D4", "D2d") and l in (1, 2):
        return _tetragonal(l)
    if group == "D2h" and l == 2:
        return _d2h()
    if group == "D3d" and l == 2:
        return _d3d()
    return None


def _d3d():
    """D3d, d shell: a1g (z^2), eg(1), eg(2) and the mixing of the two eg pairs.
    With z along the three-fold axis, the partners of an eg pair combine m = 1
    with m = -2 (and m = -1 with m = 2), which the three-fold rotation treats
    alike; eg(1) holds 1/3 of m = +-1 (the pair the cubic t2g set contributes),
    eg(2) the other 2/3. The phases e^{-+i pi/4} of the m = -+2 parts (the
    two-fold axes' orientation) and the sign of the mixing are those of the
    reference's D3d table (Core/BasicMath/BasicMath_StandardFunctions.cpp:3563),
    against which the expansion is checked."""
    l = 2
    R = real_harmonics(l)
    w = np.exp(-1j * math.pi / 4)
    a, b = math.sqrt(1.0 / 3.0), math.sqrt(2.0 / 3.0)
    e1 = [_ket(l, {1: a, -2: -b * w}), _ket(l, {-1: a, 2: b * np.conj(w)})]
    e2 = [_ket(l, {1: b, -2: a * w}), _ket(l, {-1: b, 2: -a * np.conj(w)})]
    return [_projector([R["z2"]]), _projector(e1), _projector(e2), -_mix(list(zip(e1, e2)))]


def c_matrix(k: int, q: int, l: int) -> np.ndarray:
    """<l m1|C^k_q|l m2> for m1, m2 = -l..l."""
    M = np.zeros((2 * l + 1, 2 * l + 1))
    for m2 in range(-l, l + 1):
        m1 = m2 + q
        if -l <= m1 <= l:
            M[m1 + l, m2 + l] = ck(k, l, m1, l, m2)
    return M


def expand_on_clm(V: np.ndarray, l: int, tol: float = 1e-13):
    """[[k, q, A_kq], ...] with sum A_kq <l m1|C^k_q|l m2> = V (k even <= 2l),
    coefficients below tol dropped, real ones given as floats."""
    cols, keys = [], []
    for k in range(0, 2 * l + 1, 2):
        for q in range(-k, k + 1):
            cols.append(c_matrix(k, q, l).ravel())
            keys.append((k, q))
    A = np.array(cols).T
    coef, *_ = np.linalg.lstsq(A.astype(complex), np.asarray(V, dtype=complex).ravel(), rcond=None)
    if np.max(np.abs(A @ coef - np.asarray(V).ravel())) > 1e-10 * max(1.0, np.abs(V).max()):
        raise ValueError("the operator is not an even-parity crystal field of the shell")
    out = []
    for (k, q), c in zip(keys, coef):
        if abs(c) <= tol:
            continue
        c = complex(c)
        out.append([k, q, c.real if abs(c.imag) <= tol else c])
    return out


def potential_expanded_on_clm(group: str, l: int, energies) -> list:
    """[[k, m, A_km], ...] of the crystal-field potential of `group` on the l
    shell with the given parameter energies, or KeyError when the group and l
    are not tabulated."""
    ops = _group_ops(group, l)
    if ops is None:
        raise KeyError(group)
    e = [complex(x) for x in np.atleast_1d(energies)]
    if len(e) < len(ops):
        raise ValueError(f"{group} on l = {l} takes {len(ops)} energies, got {len(e)}")
    V = sum(ei * O for ei, O in zip(e, ops))
    return expand_on_clm(V, l)
