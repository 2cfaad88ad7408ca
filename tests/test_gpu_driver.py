"""Driver semantics of Eigensystem / CreateSpectra and OperatorPsi on the GPU:
the push-form output basis of H*psi on a space far too large to enumerate,
the Hermiticity repair, Eigensystem over several start sectors and without
basis expansion, the Krylov basis of the Lanczos method, and the batched
Lanczos of several (T, psi) pairs."""
import numpy as np
import pytest
import torch

import oracle
from models import chain, dense_from_terms, jw_matrices
from quanty_amd import det as D
from quanty_amd import operators as O
from quanty_amd.api import (CreateSpectra, Eigensystem, NewOperator, NewWavefunction, Operator, dot,
                            hermitian_part)

pytestmark = pytest.mark.gpu


def _random_terms(nf, rng, n1=40, n2=12):
    """A random particle-conserving operator (not Hermitian): one- and two-body strings."""
    terms = []
    for _ in range(n1):
        i, j = rng.integers(0, nf, 2)
        terms.append(O.written([O.cr(int(i)), O.an(int(j))], complex(rng.normal(), rng.normal())))
    for _ in range(n2):
        a, b, c, d = rng.choice(nf, 4, replace=False)
        terms.append(O.written([O.cr(int(a)), O.cr(int(b)), O.an(int(c)), O.an(int(d))],
                               complex(rng.normal(), rng.normal())))
    return terms


def _oracle_vec(psi):
    w, c = psi.host()
    return oracle.Vec.from_arrays(psi.NF, D.words_to_bytes(w, psi.NF), c)


def _as_dict(nf, vec):
    d, v = vec.arrays()
    words = D.bytes_to_words(d, nf)
    return {tuple(int(x) for x in row): complex(val) for row, val in zip(words, v)}


@pytest.mark.parametrize("nf", [60, 100])
def test_hpsi_on_a_space_too_large_to_enumerate(nf):
    """H*psi of a 3-determinant psi in the nf/2-particle space of nf modes
    (C(60,30) ~ 1.2e17 determinants): the output basis is the image of psi's
    determinants (the reference's push form), equal to the oracle's OperatorPsi
    element by element; psi*H*psi and <psi|H|psi> agree with the oracle's dot."""
    rng = np.random.default_rng(nf)
    terms = _random_terms(nf, rng)
    H = Operator(nf, terms)
    half = nf // 2
    strings = []
    for _ in range(3):
        occ = rng.choice(nf, half, replace=False)
        strings.append("".join("1" if i in set(occ) else "0" for i in range(nf)))
    psi = NewWavefunction(nf, 0, [[s, complex(rng.normal(), rng.normal())] for s in strings])
    y = H * psi
    want = _as_dict(nf, oracle.Op(nf, terms).apply(_oracle_vec(psi)))
    w, c = y.host()
    got = {tuple(int(x) for x in row): complex(v) for row, v in zip(w, c)}
    assert set(got) <= set(want)
    scale = max(abs(v) for v in want.values())
    for k, v in want.items():
        assert abs(got.get(k, 0) - v) <= 1e-13 * scale
    # <psi|H|psi> two ways against the oracle
    e_ref = _oracle_vec(psi).dot(oracle.Op(nf, terms).apply(_oracle_vec(psi)))
    assert abs(dot(psi, y) - e_ref) <= 1e-12 * abs(e_ref) + 1e-14
    Hd = Operator(nf, O.dagger(terms))
    assert abs(np.conj((psi * Hd) * psi) - e_ref) <= 1e-12 * abs(e_ref) + 1e-14


def _chain_H(nbath=4, kind=0):
    nf, terms = chain(nbath, kind)
    return nf, terms, Operator(nf, terms)


def _sector_eigs(nf, terms, npart_list):
    """Eigenvalues of the dense Jordan-Wigner matrix inside the given particle numbers."""
    Hd = dense_from_terms(nf, terms, jw_matrices(nf))
    idx = [i for i in range(2 ** nf) if bin(i).count("1") in npart_list]
    return np.linalg.eigvalsh(Hd[np.ix_(idx, idx)])


@pytest.mark.parametrize("border", [1513, 1])
def test_eigensystem_repairs_an_almost_hermitian_operator(border):
    """An operator Hermitian up to Zero is replaced by its Hermitian part
    (LuaInterface/ssp.cpp:4153-4205) and solved: the energies are those of
    (H + H^+)/2, by the dense method (exact) and by the Krylov solver
    (DenseBorder 1, converged to the variance Zero = 1e-8)."""
    nf, terms, H = _chain_H(4, 1)
    pert = [O.written([O.cr(1), O.an(4)], 3e-9), O.written([O.cr(6), O.cr(0), O.an(2), O.an(9)], 2e-9j)]
    Hp = H + Operator(nf, pert)
    Hh = hermitian_part(Hp, 1e-8)
    want = _sector_eigs(nf, list(Hh.terms.items()), [5])[:4]
    psis = Eigensystem(Hp, [nf, 0, ["1" * nf, 5, 5]], 4, [["Zero", 1e-8], ["DenseBorder", border]])
    E = np.array([dot(p, Hh * p) for p in psis]).real
    tol = 1e-10 if border > 1 else 1e-7
    assert np.allclose(E, want, rtol=tol, atol=tol)


def test_eigensystem_over_two_start_sectors():
    """StartRestrictions allowing 5 or 6 electrons: the lowest states of the
    union of both sectors, ascending (the reference's BlockGroundState grows
    the space from start functions in both)."""
    nf, terms, H = _chain_H(4, 0)
    H = H + Operator(nf, O.number(list(range(nf)), [0.7] * nf))  # separates the two sectors' energies
    tot = list(O.add(terms, O.number(list(range(nf)), [0.7] * nf)))
    want = np.sort(np.concatenate([_sector_eigs(nf, tot, [5]), _sector_eigs(nf, tot, [6])]))[:6]
    psis = Eigensystem(H, [nf, 0, ["1" * nf, 5, 6]], 6, [["Zero", 1e-14]])
    E = np.array([dot(p, H * p) for p in psis]).real
    assert np.allclose(E, want, rtol=1e-10, atol=1e-10)
    assert np.all(np.diff(E) >= -1e-12)
    assert {p.particle_numbers().pop() for p in psis} == {5, 6}


def test_eigensystem_without_basis_expansion():
    """ExpandBasis = false: the determinants of the StartRestrictions are the
    basis (BlockLanczosGroundStateConserveBasis, LuaInterface/ssp.cpp:4235-4238):
    the energies are those of H projected on that set."""
    nf, terms, H = _chain_H(4, 0)
    start = [nf, 0, ["1" * nf, 5, 5], ["11" + "0" * (nf - 2), 1, 1]]
    psis = Eigensystem(H, start, 3, [["ExpandBasis", False], ["Zero", 1e-14]])
    Hd = dense_from_terms(nf, terms, jw_matrices(nf))
    idx = [i for i in range(2 ** nf) if bin(i).count("1") == 5 and bin(i >> (nf - 2)).count("1") == 1]
    want = np.linalg.eigvalsh(Hd[np.ix_(idx, idx)])[:3]
    E = np.array([dot(p, H * p) for p in psis]).real
    # H*p leaves the start set; <p|H|p> only sees its projection
    assert np.allclose(E, want, rtol=1e-10, atol=1e-10)
    assert all(p._basis.n == len(idx) for p in psis)


def test_krylov_basis_of_the_lanczos_method():
    """KrylovBasis = true with the Lanczos method
    (LanczosTriDiagonalizeKrylovCC, Core/Lanczos.cpp:3705-3854): the returned
    vectors are orthonormal and V^+ H V is the returned tridiagonal table
    (a shifted back by E0); the spectrum equals the plain method's."""
    nf, terms, H = _chain_H(5, 1)
    psi = NewWavefunction(nf, 0, [["10" + "10" * 5, 1.0], ["01" + "10" * 5, 0.5j]])
    psi = psi / np.sqrt(1.25)
    T = NewOperator("An", nf, 0)
    opts = [["Method", "Lanczos"], ["NTri", 30], ["Emin", -6.0], ["Emax", 6.0], ["NE", 300], ["Gamma", 0.2]]
    spec, tables, basis = CreateSpectra(H, T, psi, opts + [["KrylovBasis", True]])
    E0 = dot(psi, H * psi).real
    a, b = tables[0]["a"], tables[0]["b"]
    m = len(basis)
    assert 2 <= m <= 30
    G = np.array([[dot(x, y) for y in basis] for x in basis])
    assert np.max(np.abs(G - np.eye(m))) < 1e-12
    Hm = np.array([[dot(x, H * y) for y in basis] for x in basis])
    Tm = np.diag(a[1:m + 1] + E0) + np.diag(b[1:m], 1) + np.diag(b[1:m], -1)
    assert np.max(np.abs(Hm - Tm)) < 1e-10
    # the first steps, before the plain recurrence loses orthogonality, agree
    _, plain = CreateSpectra(H, T, psi, opts)
    assert np.allclose(plain[0]["a"][:10], a[:10], rtol=1e-10, atol=1e-10)
    assert np.allclose(plain[0]["b"][:10], b[:10], rtol=1e-10, atol=1e-12)


def test_batched_pairs_equal_one_at_a_time():
    """Non-tensor CreateSpectra runs its (T, psi) pairs as slots of one batch:
    the tables equal one Lanczos run per pair on the same binding."""
    from quanty_amd.api import _lanczos_many
    from quanty_amd.engine import Basis, Hop
    from quanty_amd import problems as P
    from quanty_amd import default_context
    pr = P.anderson(seed=3, nbath=2, nelec=18, full_baths=1, max_holes=2, max_elec=2)
    ctx = default_context()
    basis = Basis.enumerate(pr["NF"], pr["restrictions"], ctx)
    hop = Hop(Operator(pr["NF"], pr["H"]).device(ctx), basis, mode="assembled")
    g = np.random.default_rng(0)
    vs = [torch.from_numpy(g.normal(size=basis.n) + 1j * g.normal(size=basis.n)).cuda() for _ in range(5)]
    many = _lanczos_many(hop, vs, 30, ctx)
    for v, (a, b) in zip(vs, many):
        a1, b1 = hop.lanczos(v, 30)
        assert np.allclose(a, a1, rtol=1e-10, atol=1e-10) and np.allclose(b, b1, rtol=1e-10, atol=0)
