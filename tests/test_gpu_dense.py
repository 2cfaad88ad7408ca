"""The dense small-problem path and the determinant closure on the GPU:
- qs_basis_closure against the oracle's push-form H*psi iterated to a fixed
  point (the reference grows its dense basis this way, ssp.cpp:2259-2316),
  one- and two-word determinants, restrictions, the max_dim cut;
- qs_hop_dense against Jordan-Wigner matrices (square and rectangular);
- qs_dense_poles / qs_pole_spectrum against numpy's eigvalsh and the exact
  resolvent x^H (z - H)^-1 y, scalar and tensor;
- CreateSpectra's dense method against its Lanczos method (exhausted Krylov
  space) and against the exact resolvent when NTri is below the dimension.
Tolerances: eigenvalues 1e-12 of the spectral radius; spectra 1e-10 relative
L-infinity (dense vs resolvent) and the parity bar 1e-8 (dense vs Lanczos)."""
import math

import numpy as np
import pytest
import torch

import oracle
from models import chain, chain_states, dense_from_terms, jw_matrices
from quanty_amd import det as D
from quanty_amd import operators as O
from quanty_amd import problems as P
from quanty_amd.api import CreateSpectra, NewOperator, NewWavefunction, Operator, operator_psi
from quanty_amd.engine import Basis, Hop, Op, dense_poles, pole_spectrum

pytestmark = pytest.mark.gpu


def oracle_closure(nf, terms, seed_words, restrictions=None, seed=0):
    """Fixed point of S <- S u targets(H, S) with the oracle's push-form H*psi."""
    rng = np.random.default_rng(seed)
    op = oracle.Op(nf, terms)
    S = {tuple(int(x) for x in w) for w in np.asarray(seed_words, dtype=np.uint64).reshape(-1, D.nwords(nf))}
    frontier = sorted(S)
    while frontier:
        fw = np.array(frontier, dtype=np.uint64).reshape(-1, D.nwords(nf))
        vals = rng.normal(size=len(fw)) + 1j * rng.normal(size=len(fw))
        v = oracle.Vec.from_arrays(nf, D.words_to_bytes(fw, nf), vals)
        d, _ = op.apply(v, restrictions=restrictions).arrays()
        new = {tuple(int(x) for x in w) for w in D.bytes_to_words(d, nf)} - S
        S |= new
        frontier = sorted(new)
    return np.array(sorted(S), dtype=np.uint64).reshape(-1, D.nwords(nf))


def random_terms(nf, modes, n1, n2, rng, hermitian=True):
    terms = []
    for _ in range(n1):
        i, j = rng.choice(modes, 2)
        v = complex(*rng.normal(size=2))
        terms.append(((int(j), 0x8000 | int(i)), v))
        if hermitian:
            terms.append(((int(i), 0x8000 | int(j)), v.conjugate()))
    for _ in range(n2):
        i, j, k, l = rng.choice(modes, 4, replace=False)
        v = complex(*rng.normal(size=2))
        terms.append(((int(l), int(k), 0x8000 | int(j), 0x8000 | int(i)), v))
        if hermitian:
            terms.append(((int(i), int(j), 0x8000 | int(k), 0x8000 | int(l)), v.conjugate()))
    return terms


def det_words(nf, occ):
    x = 0
    for m in occ:
        x |= 1 << (64 * D.nwords(nf) - 1 - m)
    if D.nwords(nf) == 1:
        return np.array([[x]], dtype=np.uint64)
    return np.array([[x >> 64, x & ((1 << 64) - 1)]], dtype=np.uint64)


def test_closure_chain_matches_oracle(ctx):
    nf, terms = chain(5, 1)
    seed = np.concatenate([det_words(nf, [0, 1]), det_words(nf, [0, 4])])
    b = Basis.closure(Op(nf, terms, ctx), Basis.from_words(nf, seed, ctx))
    assert b.n == 6 * 6 + 15  # (one up, one down) and (two up) electrons on 6 sites
    assert np.array_equal(b.words(), oracle_closure(nf, terms, seed))


@pytest.mark.parametrize("kw", [dict(seed=3, nbath=1, nelec=12, full_baths=1, max_holes=3, max_elec=0,
                                     d_range=(0, 10)),
                                dict(seed=5, nbath=2, nelec=18, full_baths=1, max_holes=2, max_elec=2)])
def test_closure_with_restrictions_matches_oracle(kw, ctx):
    pr = P.anderson(**kw)
    nf = pr["NF"]
    sector = Basis.enumerate(nf, pr["restrictions"], ctx)
    rng = np.random.default_rng(1)
    seed = sector.words()[rng.choice(sector.n, 3, replace=False)]
    b = Basis.closure(Op(nf, pr["H"], ctx), Basis.from_words(nf, seed, ctx), pr["restrictions"])
    want = oracle_closure(nf, pr["H"], seed, restrictions=pr["restrictions"])
    assert np.array_equal(b.words(), want)
    # the closure lies in the sector; the cut at max_dim returns None
    assert np.all(sector.index(b.words()) >= 0)
    assert Basis.closure(Op(nf, pr["H"], ctx), Basis.from_words(nf, seed, ctx), pr["restrictions"],
                         max_dim=b.n - 1) is None
    assert Basis.closure(Op(nf, pr["H"], ctx), Basis.from_words(nf, seed, ctx), pr["restrictions"],
                         max_dim=b.n).n == b.n


def test_closure_two_word_determinants(ctx):
    """76 modes: terms couple modes of both words; the closure of one 3-electron
    determinant over 12 active modes."""
    nf = 76
    rng = np.random.default_rng(4)
    modes = np.r_[0:4, 62:66, 72:76]
    terms = random_terms(nf, modes, 10, 6, rng, hermitian=False)
    seed = det_words(nf, [0, 63, 74])
    b = Basis.closure(Op(nf, terms, ctx), Basis.from_words(nf, seed, ctx))
    want = oracle_closure(nf, terms, seed)
    assert b.n == len(want) and b.n >= 10
    assert np.array_equal(b.words(), want)
    # electrons moved between the words
    assert len(np.unique(b.words()[:, 0])) > 1 and len(np.unique(b.words()[:, 1])) > 1


def _sector_dense(nf, terms, rows, cols):
    """Jordan-Wigner matrix restricted to the row and column bases."""
    Hf = dense_from_terms(nf, terms, jw_matrices(nf))
    shift = np.uint64(64 - nf)
    ri = (rows.words()[:, 0] >> shift).astype(np.int64)
    ci = (cols.words()[:, 0] >> shift).astype(np.int64)
    return Hf[np.ix_(ri, ci)]


def test_dense_matrix_matches_jordan_wigner(ctx):
    nf = 8
    rng = np.random.default_rng(2)
    terms = random_terms(nf, np.arange(nf), 12, 10, rng, hermitian=False)
    terms.append(((0x8000 | 3, 3), 0.7 - 0.1j))  # a diagonal term
    rows = Basis.enumerate(nf, [("11111111", 4, 4)], ctx)
    hop = Hop(Op(nf, terms, ctx), rows, mode="matfree")
    M = hop.dense().cpu().numpy()
    assert np.allclose(M, _sector_dense(nf, terms, rows, rows), atol=1e-14, rtol=0)
    # rectangular: a creation-type operator from 3 to 4 electrons
    cols = Basis.enumerate(nf, [("11111111", 3, 3)], ctx)
    tterms = [((0x8000 | i,), complex(*rng.normal(size=2))) for i in range(nf)]
    M2 = Hop(Op(nf, tterms, ctx), rows, cols, mode="matfree").dense().cpu().numpy()
    assert M2.shape == (rows.n, cols.n)
    assert np.allclose(M2, _sector_dense(nf, tterms, rows, cols), atol=1e-14, rtol=0)


def test_dense_poles_match_numpy_resolvent(ctx):
    nf = 8
    rng = np.random.default_rng(3)
    terms = random_terms(nf, np.arange(nf), 14, 12, rng)
    rows = Basis.enumerate(nf, [("11111111", 4, 4)], ctx)
    Href = _sector_dense(nf, terms, rows, rows)
    assert np.allclose(Href, Href.conj().T, atol=1e-13)
    Hd = Hop(Op(nf, terms, ctx), rows, mode="matfree").dense()
    X = torch.from_numpy(rng.normal(size=(3, rows.n)) + 1j * rng.normal(size=(3, rows.n))).to(Hd.device)
    E, Pm = dense_poles(Hd, X, ctx)
    Eref = np.linalg.eigvalsh(Href)
    scale = np.max(np.abs(Eref))
    assert np.max(np.abs(E - Eref)) < 1e-12 * scale
    Z = Hd.cpu().numpy()
    assert np.max(np.abs(Href @ Z - Z * E[None, :])) < 1e-11 * scale
    # spectra from the poles against the exact resolvent
    E0, emin, emax, ne, gamma = 0.3, -12.0, 12.0, 240, 0.35
    G = pole_spectrum(E, Pm, E0, emin, emax, ne, gamma, ctx=ctx)
    Gt = pole_spectrum(E, Pm, E0, emin, emax, ne, gamma, tensor=True, ctx=ctx)
    Xh = X.cpu().numpy()
    for i in range(0, ne + 1, 7):
        w = emin + i * (emax - emin) / ne
        R = np.linalg.solve((w + E0 + 0.5j * gamma) * np.eye(rows.n) - Href, Xh.T)
        full = Xh.conj() @ R  # [r, c] = x_r^H (z - H)^-1 x_c
        assert np.allclose(G[:, i], np.diag(full), rtol=1e-10, atol=1e-12 * np.max(np.abs(full)))
        wt = (i / ne) * (emax - emin) + emin
        Rt = np.linalg.solve((wt + E0 + 0.5j * gamma) * np.eye(rows.n) - Href, Xh.T)
        assert np.allclose(Gt[i], Xh.conj() @ Rt, rtol=1e-10, atol=1e-12 * np.max(np.abs(full)))


def _chain_api(nbath, kind):
    nf, terms = chain(nbath, kind)
    return nf, Operator(nf, dict(terms))


def test_create_spectra_dense_equals_lanczos_when_exhausted(ctx):
    nf, H = _chain_api(5, 1)
    T = [NewOperator("Cr", nf, 0), NewOperator("Cr", nf, 1)]
    strings, coefs = chain_states(nf, 2)
    psi = NewWavefunction(nf, 0, list(zip(strings, coefs)))
    opts = [["Emin", -6.0], ["Emax", 8.0], ["NE", 400], ["Gamma", 0.3], ["NTri", 60]]
    sd, td = CreateSpectra(H, T, psi, opts)
    sl, tl = CreateSpectra(H, T, psi, opts + [["Method", "Lanczos"]])
    assert td[0]["type"] == "ListOfPoles" and tl[0]["type"] == "Tri"
    for i in range(2):
        assert np.max(np.abs(sd.G[i] - sl.G[i])) <= 1e-8 * np.max(np.abs(sl.G[i]))
    # total weight of the poles = |T psi|^2; the table evaluates to the spectrum
    tp = operator_psi(T[0], psi)
    assert abs(td[0]["b"].sum() - np.sum(np.abs(tp.host()[1]) ** 2)) < 1e-12
    w = -6.0 + 37 * 14.0 / 400
    assert abs(td[0](w, 0.3) - sd.G[0][37]) < 1e-10 * np.max(np.abs(sd.G[0]))
    # tensor: dense equals block Lanczos
    st, _ = CreateSpectra(H, T, psi, opts + [["Tensor", True]])
    sb, _ = CreateSpectra(H, T, psi, opts + [["Tensor", True], ["Method", "Lanczos"]])
    assert np.max(np.abs(st.G - sb.G)) <= 1e-8 * np.max(np.abs(sb.G))
    # KrylovBasis: the eigenstates of H on the closure
    out = CreateSpectra(H, T[0], psi, opts + [["KrylovBasis", True]])
    assert len(out) == 3
    zs = out[2]
    cb = zs[0]._basis
    assert len(zs) == cb.n
    hop = Hop(H.device(ctx), cb, mode="matfree")
    for z in zs:
        y = hop.apply(z._vec.contiguous())
        lam = torch.vdot(z._vec, y).item()
        assert abs(lam.imag) < 1e-12 * max(1.0, abs(lam))
        assert torch.linalg.norm(y - lam * z._vec).item() < 1e-11 * max(1.0, abs(lam))


def test_create_spectra_dense_is_exact_below_ntri(ctx):
    """NTri smaller than the dimension: the dense method gives the exact resolvent
    (the reference's default for small spaces), Lanczos only approximates it."""
    nf, terms = chain(3, 0)
    H = Operator(nf, dict(terms))
    T = NewOperator("Cr", nf, 0)
    strings, coefs = chain_states(nf, 1)
    psi = NewWavefunction(nf, 0, list(zip(strings, coefs)))
    opts = [["Emin", -6.0], ["Emax", 8.0], ["NE", 200], ["Gamma", 0.2], ["NTri", 6], ["E0", 0.0]]
    sd, _ = CreateSpectra(H, T, psi, opts)
    tp = operator_psi(T, psi)
    w, c = tp.host()
    Hf = dense_from_terms(nf, terms, jw_matrices(nf))
    idx = (w[:, 0] >> np.uint64(64 - nf)).astype(np.int64)
    x = np.zeros(2 ** nf, dtype=complex)
    x[idx] = c
    for i in range(0, 201, 25):
        om = -6.0 + i * 14.0 / 200
        g = np.vdot(x, np.linalg.solve((om + 0.1j) * np.eye(2 ** nf) - Hf, x))
        assert abs(sd.G[0][i] - g) <= 1e-10 * np.max(np.abs(sd.G[0]))


def test_dense_pole_range_is_the_closure_spectrum(ctx):
    """Default Emin/Emax come from the eigenvalues of H on the closure of T psi,
    not on the whole particle sector. Two orbitals a, b with up (modes 0, 1) and
    down (2, 3) spins: hopping, U n_up n_dn on each orbital, spin exchange
    (couples all four modes, so they form one conserved group) and a Zeeman
    field. H keeps N_up and N_dn, so from one up and one down electron the
    closure is the 4 states with Sz = 0, while the two-electron sector of the
    group also holds the up-up and down-down states at -+2h."""
    nf, t, U, J, h = 4, 1.0, 5.0, 0.7, 3.0
    terms = [((1, 0x8000 | 0), t), ((0, 0x8000 | 1), t), ((3, 0x8000 | 2), t), ((2, 0x8000 | 3), t),
             ((0, 0x8000 | 0, 2, 0x8000 | 2), U), ((1, 0x8000 | 1, 3, 0x8000 | 3), U),
             ((1, 2, 0x8000 | 3, 0x8000 | 0), J), ((3, 0, 0x8000 | 1, 0x8000 | 2), J),
             ((0, 0x8000 | 0), h), ((1, 0x8000 | 1), h), ((2, 0x8000 | 2), -h), ((3, 0x8000 | 3), -h)]
    H = Operator(nf, dict(terms))
    psi = NewWavefunction(nf, 0, [("1000", 1.0)])
    T = NewOperator("Cr", nf, 2)
    spec, tables = CreateSpectra(H, T, psi, [["NE", 100], ["E0", 0.0]])
    tp = operator_psi(T, psi)
    cb = Basis.closure(Op(nf, terms), Basis.from_words(nf, tp.host()[0]))
    assert cb.n == 4
    Hc = Hop(Op(nf, terms), cb, mode="matfree").dense().cpu().numpy()
    ev = np.linalg.eigvalsh(Hc)
    assert abs(spec.EminPole - ev[0]) < 1e-12 * np.max(np.abs(ev))
    assert abs(spec.EmaxPole - ev[-1]) < 1e-12 * np.max(np.abs(ev))
    full = np.linalg.eigvalsh(dense_from_terms(nf, terms, jw_matrices(nf)))
    assert full[0] < ev[0] - 1 and full[-1] > ev[-1] + 1  # the whole space reaches further


def test_eigensystem_dense_equals_krylov(ctx):
    """Eigensystem below DenseBorder diagonalizes the determinant matrix; above
    it runs block Lanczos. Both give the same energies (1e-10) and eigenstates
    with residuals at rounding level."""
    from quanty_amd.api import Eigensystem, _expect
    nf, terms = chain(4, 0)
    H = Operator(nf, dict(terms))
    start = [nf, 0, ["10" * (nf // 2), 2, 2], ["01" * (nf // 2), 2, 2]]
    dense = Eigensystem(H, start, 4)
    kry = Eigensystem(H, start, 4, [["DenseBorder", 0], ["Zero", 1e-14]])
    assert dense[0]._basis.n == 100 and dense[0]._basis.n <= 1513
    Ed = [_expect(H, p).real for p in dense]
    Ek = [_expect(H, p).real for p in kry]
    assert np.allclose(Ed, Ek, rtol=1e-10, atol=1e-12)
    assert np.all(np.diff(Ed) >= -1e-12)
    hop = Hop(H.device(ctx), dense[0]._basis, mode="matfree")
    for p, e in zip(dense, Ed):
        y = hop.apply(p._vec.contiguous())
        assert torch.linalg.norm(y - e * p._vec).item() < 1e-11 * max(1.0, abs(e))


@pytest.mark.parametrize("degenerate", [False, True])
def test_dense_poles_with_vectors_equals_projection_path(degenerate, ctx):
    """qs_dense_poles with all eigenvectors (tridiagonal divide and conquer, then
    the back-transform Q Z) against the projections-only path with the host QL
    and with the GPU divide and conquer, and numpy: the same eigenvalues and
    spectra (degenerate poles may split their weight differently; the spectra
    see only the sums); the returned eigenvectors orthonormal with H z = E z and
    P = Z^+ X."""
    from quanty_amd._lib import QS_PARAM_TRIDIAG_DC
    n = 300
    rng = np.random.default_rng(11 + degenerate)
    if degenerate:
        A = rng.normal(size=(n // 3, n // 3)) + 1j * rng.normal(size=(n // 3, n // 3))
        M = np.kron(np.eye(3), A + A.conj().T)
    else:
        B = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        M = B + B.conj().T
    Xh = rng.normal(size=(4, n)) + 1j * rng.normal(size=(4, n))
    X = torch.from_numpy(Xh).cuda()
    out = {}
    try:
        for key, vec, dc in (("vectors", True, 256), ("ql", False, 10 ** 9), ("dc", False, 1)):
            ctx.set_param(QS_PARAM_TRIDIAG_DC, dc)
            T = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()
            out[key] = dense_poles(T.t(), X, ctx, vectors=vec)
            if vec:
                Z = T.t().cpu().numpy()
    finally:
        ctx.set_param(QS_PARAM_TRIDIAG_DC, 256)
    E1, P1 = out["vectors"]
    scale = np.max(np.abs(E1))
    assert np.max(np.abs(E1 - np.linalg.eigvalsh(M))) < 1e-12 * scale
    assert np.all(np.isfinite(Z))
    assert np.max(np.abs(M @ Z - Z * E1[None, :])) < 1e-11 * scale
    assert np.max(np.abs(Z.conj().T @ Z - np.eye(n))) < 1e-12
    assert np.max(np.abs(P1 - Z.conj().T @ Xh.T)) < 1e-12 * np.max(np.abs(Xh)) * np.sqrt(n)
    G1 = pole_spectrum(E1, P1, 0.0, -40.0, 40.0, 400, 0.5, ctx=ctx)
    Gt1 = pole_spectrum(E1, P1, 0.0, -40.0, 40.0, 400, 0.5, tensor=True, ctx=ctx)
    for key in ("ql", "dc"):
        E0, P0 = out[key]
        assert np.max(np.abs(E0 - E1)) < 1e-12 * scale
        assert np.allclose((np.abs(P0) ** 2).sum(0), (np.abs(Xh) ** 2).sum(1), rtol=1e-12)
        G0 = pole_spectrum(E0, P0, 0.0, -40.0, 40.0, 400, 0.5, ctx=ctx)
        assert np.max(np.abs(G0 - G1)) < 1e-10 * np.max(np.abs(G1))
        Gt0 = pole_spectrum(E0, P0, 0.0, -40.0, 40.0, 400, 0.5, tensor=True, ctx=ctx)
        assert np.max(np.abs(Gt0 - Gt1)) < 1e-10 * np.max(np.abs(Gt1))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 300])
@pytest.mark.parametrize("kind", ["complex", "real", "blocks", "diagonal"])
def test_dense_tridiagonal_divide_and_conquer_equals_host_ql(n, kind, ctx):
    """The tridiagonal eigenproblem by rocSOLVER's divide and conquer on the GPU
    against the host QL, on the same reduction: eigenvalues to 1e-12, weight
    sums conserved, spectra (plain and tensor) to 1e-10; all eigenvectors from
    either route orthonormal with H z = E z (split matrices included: the
    block and diagonal kinds reduce to tridiagonal matrices with zero
    couplings)."""
    from quanty_amd._lib import QS_PARAM_TRIDIAG_DC
    rng = np.random.default_rng(7000 + n + len(kind))
    M = _hermitian(n, rng, kind)
    Xh = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    X = torch.from_numpy(Xh).cuda()
    out = {}
    try:
        for dc in (1, 10 ** 9):
            ctx.set_param(QS_PARAM_TRIDIAG_DC, dc)
            T = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()
            out[dc] = dense_poles(T.t(), X, ctx, vectors=False)
    finally:
        ctx.set_param(QS_PARAM_TRIDIAG_DC, 256)
    (E1, P1), (E0, P0) = out[1], out[10 ** 9]
    scale = max(1.0, np.max(np.abs(E0)))
    assert np.max(np.abs(E1 - np.linalg.eigvalsh(M))) < 1e-12 * scale
    assert np.max(np.abs(E1 - E0)) < 1e-12 * scale
    assert np.allclose((np.abs(P1) ** 2).sum(0), (np.abs(Xh) ** 2).sum(1), rtol=1e-12)
    G1 = pole_spectrum(E1, P1, 0.0, -40.0, 40.0, 300, 0.5, ctx=ctx)
    G0 = pole_spectrum(E0, P0, 0.0, -40.0, 40.0, 300, 0.5, ctx=ctx)
    assert np.max(np.abs(G1 - G0)) < 1e-10 * np.max(np.abs(G0))
    Gt1 = pole_spectrum(E1, P1, 0.0, -40.0, 40.0, 300, 0.5, tensor=True, ctx=ctx)
    Gt0 = pole_spectrum(E0, P0, 0.0, -40.0, 40.0, 300, 0.5, tensor=True, ctx=ctx)
    assert np.max(np.abs(Gt1 - Gt0)) < 1e-10 * np.max(np.abs(Gt0))
    T = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()
    E, _ = dense_poles(T.t(), None, ctx, vectors=True)
    Z = T.t().cpu().numpy()
    assert np.max(np.abs(E - E0)) < 1e-12 * scale
    assert np.max(np.abs(M @ Z - Z * E[None, :])) < 1e-11 * scale
    assert np.max(np.abs(Z.conj().T @ Z - np.eye(n))) < 1e-12


@pytest.mark.parametrize("n", [100, 400])
def test_dense_poles_degenerate_spectrum(n, ctx):
    """Exactly degenerate eigenvalues (two copies of one block): every eigenvector
    finite, H z = E z and Z orthonormal at rounding level."""
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n // 2, n // 2)) + 1j * rng.normal(size=(n // 2, n // 2))
    M = np.kron(np.eye(2), A + A.conj().T)
    T = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()
    E, _ = dense_poles(T.t(), None, ctx)
    Z = T.t().cpu().numpy()
    scale = np.max(np.abs(E))
    assert np.all(np.isfinite(Z))
    assert np.max(np.abs(E - np.linalg.eigvalsh(M))) < 1e-12 * scale
    assert np.max(np.abs(M @ Z - Z * E[None, :])) < 1e-11 * scale
    assert np.max(np.abs(Z.conj().T @ Z - np.eye(n))) < 1e-11


def test_dense_two_word_determinants_match_matrix_free(ctx):
    """76 modes (two-word determinants): every column of the dense matrix equals
    the matrix-free H e_j, and CreateSpectra's dense method equals its Lanczos
    method when the Krylov space is exhausted."""
    nf = 76
    rng = np.random.default_rng(8)
    modes = np.r_[0:4, 62:66, 72:76]
    terms = random_terms(nf, modes, 10, 6, rng)
    op = Op(nf, terms, ctx)
    cb = Basis.closure(op, Basis.from_words(nf, det_words(nf, [1, 63, 73]), ctx))
    n = cb.n
    M = Hop(op, cb, mode="matfree").dense().cpu().numpy()
    mf = Hop(op, cb, mode="matfree")
    for j in range(0, n, max(1, n // 12)):
        e = torch.zeros(n, dtype=torch.complex128, device="cuda")
        e[j] = 1
        col = mf.apply(e).cpu().numpy()
        assert np.allclose(M[:, j], col, atol=1e-14, rtol=0)
    assert np.allclose(M, M.conj().T, atol=1e-13)
    H = Operator(nf, dict(terms))
    psi = NewWavefunction(nf, 0, [("".join("1" if i in (1, 63) else "0" for i in range(nf)), 1.0)])
    T = NewOperator("Cr", nf, 73)
    opts = [["Emin", -8.0], ["Emax", 8.0], ["NE", 300], ["Gamma", 0.25], ["NTri", 2 * n + 10], ["E0", 0.0]]
    sd, _ = CreateSpectra(H, T, psi, opts)
    sl, _ = CreateSpectra(H, T, psi, opts + [["Method", "Lanczos"]])
    assert np.max(np.abs(sd.G - sl.G)) <= 1e-8 * np.max(np.abs(sl.G))


@pytest.mark.parametrize("degenerate", [False, True])
def test_dense_lowest_eigenvectors(degenerate, ctx):
    """qs_dense_lowest: every eigenvalue, and the k lowest eigenvectors
    orthonormal with H z = E z at rounding level (degenerate included), with
    the tridiagonal eigenvectors from the GPU divide and conquer and from the
    host QL."""
    from quanty_amd.engine import dense_lowest
    n, k = 240, 7
    rng = np.random.default_rng(21 + degenerate)
    if degenerate:
        A = rng.normal(size=(n // 2, n // 2)) + 1j * rng.normal(size=(n // 2, n // 2))
        M = np.kron(np.eye(2), A + A.conj().T)
    else:
        B = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        M = B + B.conj().T
    from quanty_amd._lib import QS_PARAM_TRIDIAG_DC
    try:
        for dc in (1, 10 ** 9):  # divide and conquer on the GPU, host QL
            ctx.set_param(QS_PARAM_TRIDIAG_DC, dc)
            T = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()
            E, V = dense_lowest(T.t(), k, ctx)
            Z = V.cpu().numpy().T  # columns z_j
            scale = np.max(np.abs(E))
            assert np.max(np.abs(E - np.linalg.eigvalsh(M))) < 1e-12 * scale
            assert np.max(np.abs(M @ Z - Z * E[None, :k])) < 1e-11 * scale
            assert np.max(np.abs(Z.conj().T @ Z - np.eye(k))) < 1e-12
    finally:
        ctx.set_param(QS_PARAM_TRIDIAG_DC, 256)


def test_create_spectra_with_annihilated_psi(ctx):
    """A transition operator that annihilates psi gives a zero spectrum and a
    zero table, alone and beside a non-zero one, by both methods and in the
    tensor form (the reference's Lanczos on an empty wavefunction has b[0] = 0)."""
    nf, terms = chain(3, 0)
    H = Operator(nf, dict(terms))
    psi = NewWavefunction(nf, 0, [("10" + "00" * 3, 1.0)])
    kill = NewOperator("Cr", nf, 0)   # mode 0 is occupied
    live = NewOperator("Cr", nf, 1)
    opts = [["Emin", -6.0], ["Emax", 8.0], ["NE", 100], ["Gamma", 0.3], ["NTri", 20], ["E0", 0.0]]
    for method in ("Auto", "Lanczos"):
        s0, t0 = CreateSpectra(H, kill, psi, opts + [["Method", method]])
        assert np.all(s0.G == 0) and np.all(np.isfinite(s0.G))
        assert np.all(t0[0]["b"] == 0)
        s2, _ = CreateSpectra(H, [kill, live], psi, opts + [["Method", method]])
        s1, _ = CreateSpectra(H, live, psi, opts + [["Method", method]])
        assert np.all(s2.G[0] == 0)
        assert np.max(np.abs(s2.G[1] - s1.G[0])) <= 1e-12 * np.max(np.abs(s1.G[0]))
    st, _ = CreateSpectra(H, kill, psi, opts + [["Tensor", True]])
    assert np.all(st.G == 0)


def test_tensor_lanczos_default_range_from_block_eigenvalues(ctx):
    """Tensor Lanczos without Emin/Emax takes the range from the eigenvalues of
    its block tridiagonal matrix. With the block Krylov space exhausted those
    are eigenvalues of H on the closure: the pole range lies inside the dense
    method's (which sees every closure eigenvalue) and equals it when the
    Krylov space is the whole closure."""
    nf, H = _chain_api(5, 1)
    T = [NewOperator("Cr", nf, 0), NewOperator("Cr", nf, 1)]
    strings, coefs = chain_states(nf, 2)
    psi = NewWavefunction(nf, 0, list(zip(strings, coefs)))
    opts = [["NE", 400], ["Gamma", 0.3], ["NTri", 60], ["Tensor", True]]
    sl, tab = CreateSpectra(H, T, psi, opts + [["Method", "Lanczos"]])
    sd, _ = CreateSpectra(H, T, psi, opts)
    kept = tab["kept"]
    krylov_dim = int(sum(kept[1:]))
    assert sl.EminPole >= sd.EminPole - 1e-9 and sl.EmaxPole <= sd.EmaxPole + 1e-9
    tp = [operator_psi(t, psi) for t in T]
    seed = np.unique(np.concatenate([t.host()[0] for t in tp]), axis=0)
    cb = Basis.closure(Op(nf, list(H.terms.items()), ctx), Basis.from_words(nf, seed, ctx))
    if krylov_dim == cb.n:
        assert abs(sl.EminPole - sd.EminPole) < 1e-9 and abs(sl.EmaxPole - sd.EmaxPole) < 1e-9
    # the range follows the reference's formula around the pole range
    span = sl.EmaxPole - sl.EminPole
    assert abs(sl.Emin - (sl.EminPole - 0.125 * span)) < 1e-12 * max(1, abs(sl.Emin))
    assert abs(sl.Emax - (sl.EmaxPole + 0.125 * span)) < 1e-12 * max(1, abs(sl.Emax))


def test_dense_single_determinant_closure(ctx):
    """H diagonal on T psi's only determinant: a closure of one determinant,
    one pole at its energy with the whole weight."""
    nf = 6
    H = NewOperator(nf, 0, [[0, -0.0, 1.5], [1, -1, -0.25]])  # 1.5 n0 - 0.25 n1
    psi = NewWavefunction(nf, 0, [("010000", 1.0)])
    T = NewOperator("Cr", nf, 0)
    spec, tables = CreateSpectra(H, T, psi, [["Emin", -2.0], ["Emax", 4.0], ["NE", 60], ["Gamma", 0.2],
                                              ["E0", 0.0]])
    assert tables[0]["type"] == "ListOfPoles"
    assert np.allclose(tables[0]["a"], [0.0, 1.25]) and np.allclose(tables[0]["b"], [1.0])
    w = -2.0 + np.arange(61) * 0.1
    assert np.allclose(spec.G[0], 1.0 / (w - 1.25 + 0.1j), rtol=1e-12, atol=1e-14)


def test_closure_two_word_with_restrictions(ctx):
    """A restriction whose mask spans both words of 76-mode determinants cuts the
    closure exactly as the oracle's restricted push form does."""
    nf = 76
    rng = np.random.default_rng(5)
    modes = np.r_[0:4, 60:68, 72:76]
    terms = random_terms(nf, modes, 12, 6, rng, hermitian=False)
    mask = "".join("1" if i in (1, 2, 62, 63, 66) else "0" for i in range(nf))
    restr = [(mask, 1, 2)]
    seed = det_words(nf, [1, 63, 73])
    b = Basis.closure(Op(nf, terms, ctx), Basis.from_words(nf, seed, ctx), restr)
    want = oracle_closure(nf, terms, seed, restrictions=restr)
    assert b.n == len(want) and np.array_equal(b.words(), want)


@pytest.mark.parametrize("cplx", [False, True])
def test_create_spectra_pole_range_tensor(cplx):
    """TestScripts/CreateSpectraPoleRange.Quanty: a one-body H from the 20 x 20
    matrix H_ij = cos(i+2j) + cos(2i+j) (+ i sin(i-j)) + i delta_ij (1-based),
    T = c+_0, c+_1, c+_2 on the vacuum, tensor block Lanczos (DenseBorder 0),
    explicit Emin/Emax, NTri 5. The chain must not deflate (15 x 15 block
    tridiagonal matrix) and EminPole/EmaxPole are its extreme eigenvalues (1e-10),
    which by interlacing lie inside H's one-particle spectrum."""
    N = 20
    i = np.arange(1, N + 1)[:, None]
    j = np.arange(1, N + 1)[None, :]
    M = np.cos(i + 2 * j) + np.cos(2 * i + j) + np.diag(np.arange(1, N + 1)).astype(float)
    if cplx:
        M = M + 1j * np.sin(i - j)
    H = Operator(N, O.one_body(list(range(N)), list(range(N)), M))
    T = [NewOperator(N, 0, [[k, 1]]) for k in range(3)]
    psi0 = NewWavefunction(N, 0, [["0" * N, 1]])
    S, tab = CreateSpectra(H, T, psi0, [["Tensor", True], ["DenseBorder", 0], ["Emin", -30], ["Emax", 30],
                                         ["NE", 100], ["Gamma", 0.1], ["NTri", 5]])
    A, B, kept = tab["A"], tab["B"], tab["kept"]
    sizes = [int(k) for k in kept[1:] if k > 0]
    assert sum(sizes) == 15 and S.Emin == -30 and S.Emax == 30
    off = np.concatenate([[0], np.cumsum(sizes)])
    Tm = np.zeros((15, 15), dtype=complex)
    for l, p in enumerate(sizes, 1):
        Tm[off[l - 1]:off[l], off[l - 1]:off[l]] = A[l][:p, :p]
        if l < len(sizes):
            q = sizes[l]
            Tm[off[l - 1]:off[l], off[l]:off[l] + q] = B[l][:p, :q]
            Tm[off[l]:off[l] + q, off[l - 1]:off[l]] = B[l][:p, :q].conj().T
    val = np.linalg.eigvalsh(Tm)
    assert abs(S.EminPole - val[0]) < 1e-10 and abs(S.EmaxPole - val[-1]) < 1e-10
    lam = np.linalg.eigvalsh(M)
    assert lam[0] - 1e-10 <= S.EminPole and S.EmaxPole <= lam[-1] + 1e-10


def _deflation_cases(cplx):
    """The three one-body models of TestScripts/CreateSpectraDeflation.Quanty
    (1-based matrix indices as there): (N, H matrix, T mode list)."""
    def ph(k):
        return complex(math.cos(k), math.sin(k)) if cplx else 1.0

    def hop(M, i, j, t):
        M[i - 1, j - 1] += t
        M[j - 1, i - 1] += np.conj(t)

    M = np.diag([-1, 1, 0.5, -0.5, 2, -2]).astype(complex)
    for i, j, t in ((1, 3, 0.7 * ph(1)), (2, 4, 0.9 * ph(2)), (3, 5, 0.4), (4, 6, 0.6), (1, 2, 0.3 * ph(3))):
        hop(M, i, j, t)
    yield 6, M, [0, 0, 1, 1, 1]
    M = np.diag([0, 0, 0, 0, 0, 1.5, -1.5, 3, -3]).astype(complex)
    for i, j, t in ((1, 6, ph(1)), (2, 6, ph(2)), (3, 7, ph(3)), (4, 7, ph(4)), (5, 7, ph(5)), (6, 8, 0.8),
                    (7, 9, 0.5)):
        hop(M, i, j, t)
    yield 9, M, [0, 1, 2, 3, 4]
    M = np.diag([0.6, 0.6, -0.4, -0.4, -0.4, -3, -3, -3]).astype(complex)
    for k in range(3):
        hop(M, 3 + k, 6 + k, 1.2 * ph(k + 1))
    yield 8, M, [0, 1, 2, 3, 4]


@pytest.mark.parametrize("cplx", [False, True])
def test_create_spectra_block_deflation(cplx):
    """TestScripts/CreateSpectraDeflation.Quanty: tensor spectra from block
    Lanczos (DenseBorder 0, NTri 10) equal the dense method's to 1e-10 of their
    maximum when the start block is linearly dependent ({a, a, b, b, b}), when
    the chain deflates in its second block, and when the eg orbitals decouple
    after the first block (then the Krylov blocks hold 5 and 3 vectors)."""
    for case, (N, M, modes) in enumerate(_deflation_cases(cplx)):
        H = Operator(N, O.one_body(list(range(N)), list(range(N)), M))
        T = [NewOperator(N, 0, [[k, 1]]) for k in modes]
        psi0 = NewWavefunction(N, 0, [["0" * N, 1]])
        opts = [["Tensor", True], ["Emin", -10], ["Emax", 10], ["NE", 200], ["Gamma", 0.2]]
        D, _ = CreateSpectra(H, T, psi0, opts + [["DenseBorder", 100000]])
        S, tab = CreateSpectra(H, T, psi0, opts + [["DenseBorder", 0], ["NTri", 10]])
        d = np.max(np.abs(S.G - D.G)) / np.max(np.abs(D.G))
        assert d < 1e-10, (case, d)
        if case == 2:
            assert [int(k) for k in tab["kept"][1:] if k > 0] == [5, 3]


def _hermitian(n, rng, kind):
    if kind == "complex":
        B = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        return B + B.conj().T
    if kind == "real":
        B = rng.normal(size=(n, n))
        return (B + B.T).astype(complex)
    if kind == "blocks":  # reflectors with tau = 0 at every block edge
        m = max(1, n // 4)
        A = rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))
        M = np.zeros((n, n), dtype=complex)
        for b in range(0, n, m):
            e = min(n, b + m)
            M[b:e, b:e] = (A + A.conj().T)[: e - b, : e - b]
        return M
    if kind == "diagonal":  # nothing to reduce: every tau is 0
        return np.diag(rng.normal(size=n)).astype(complex)
    raise ValueError(kind)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 64, 257, 800, 1536])
@pytest.mark.parametrize("kind", ["complex", "real", "blocks", "diagonal"])
def test_dense_reduction_cooperative_equals_rocsolver(n, kind, ctx):
    """The dense method's own reduction (one cooperative launch, the matrix in
    registers, Q^+ X formed on the way; 1 to 256 workgroups) against
    rocSOLVER's zhetrd + Q^+ X and numpy: eigenvalues to 1e-12, the weight sums
    |<z_k|x>|^2 conserved, spectra to 1e-10."""
    from quanty_amd._lib import QS_PARAM_HETRD
    rng = np.random.default_rng(1000 * n + len(kind))
    M = _hermitian(n, rng, kind)
    Xh = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    X = torch.from_numpy(Xh).cuda()
    out = {}
    try:
        for mode in (1, 0):
            ctx.set_param(QS_PARAM_HETRD, mode)
            T = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()
            out[mode] = dense_poles(T.t(), X, ctx, vectors=False)
    finally:
        ctx.set_param(QS_PARAM_HETRD, 1)
    (E1, P1), (E0, P0) = out[1], out[0]
    scale = max(1.0, np.max(np.abs(E0)))
    assert np.max(np.abs(E1 - np.linalg.eigvalsh(M))) < 1e-12 * scale
    assert np.max(np.abs(E1 - E0)) < 1e-12 * scale
    assert np.allclose((np.abs(P1) ** 2).sum(0), (np.abs(Xh) ** 2).sum(1), rtol=1e-12)
    G1 = pole_spectrum(E1, P1, 0.0, -40.0, 40.0, 300, 0.5, ctx=ctx)
    G0 = pole_spectrum(E0, P0, 0.0, -40.0, 40.0, 300, 0.5, ctx=ctx)
    assert np.max(np.abs(G1 - G0)) < 1e-10 * np.max(np.abs(G0))
    Gt1 = pole_spectrum(E1, P1, 0.0, -40.0, 40.0, 300, 0.5, tensor=True, ctx=ctx)
    Gt0 = pole_spectrum(E0, P0, 0.0, -40.0, 40.0, 300, 0.5, tensor=True, ctx=ctx)
    assert np.max(np.abs(Gt1 - Gt0)) < 1e-10 * np.max(np.abs(Gt0))


def test_dense_reduction_wait_bound_ends_the_launch(ctx):
    """Every wait of the cooperative reduction is bounded. With
    QS_PARAM_WAIT_SPINS = 0 a wait gives up at its first poll that finds no
    data (at n = 800, 134 workgroups over 799 columns, some poll always does):
    that workgroup raises the shared error flag, every workgroup leaves the
    launch at its next check, and the call raises the error naming the wait.
    With the bound restored the next call is correct (the exchange buffers are
    reset per call)."""
    from quanty_amd._lib import QS_PARAM_WAIT_SPINS, QSError
    n = 800
    rng = np.random.default_rng(4242)
    M = _hermitian(n, rng, "complex")
    X = torch.from_numpy(rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))).cuda()
    default = ctx.get_param(QS_PARAM_WAIT_SPINS)
    try:
        ctx.set_param(QS_PARAM_WAIT_SPINS, 0)
        T = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()
        with pytest.raises(QSError, match="waited too long"):
            dense_poles(T.t(), X, ctx, vectors=False)
    finally:
        ctx.set_param(QS_PARAM_WAIT_SPINS, default)
    assert default == 1 << 22
    T = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()
    E, P = dense_poles(T.t(), X, ctx, vectors=False)
    assert np.max(np.abs(E - np.linalg.eigvalsh(M))) < 1e-12 * np.max(np.abs(E))
    assert np.allclose((np.abs(P) ** 2).sum(0), (np.abs(X.cpu().numpy()) ** 2).sum(1), rtol=1e-12)


@pytest.mark.parametrize("n,kind", [(2, "complex"), (130, "blocks"), (800, "complex"), (1536, "complex"),
                                    (1537, "complex")])
def test_dense_lowest_after_cooperative_reduction(n, kind, ctx):
    """qs_dense_lowest through the cooperative reduction (n <= 1536; rocSOLVER
    above): the stored reflectors back-transform the tridiagonal eigenvectors
    into orthonormal eigenvectors of H."""
    from quanty_amd.engine import dense_lowest
    rng = np.random.default_rng(n)
    M = _hermitian(n, rng, kind)
    k = min(n, 4)
    T = torch.from_numpy(np.ascontiguousarray(M.T)).cuda()
    E, V = dense_lowest(T.t(), k, ctx)
    Z = V.cpu().numpy().T
    scale = np.max(np.abs(E))
    assert np.max(np.abs(E - np.linalg.eigvalsh(M))) < 1e-11 * scale
    assert np.max(np.abs(M @ Z - Z * E[None, :k])) < 1e-10 * scale
    assert np.max(np.abs(Z.conj().T @ Z - np.eye(k))) < 1e-11
