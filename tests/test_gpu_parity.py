"""HIP path (libqs) against the oracle: bases bit-exact, H*psi to 1e-13,
Lanczos coefficients to 1e-10, continued fractions to 1e-13, golden tables.
All calls go through the C ABI."""
import math

import numpy as np
import pytest
import torch

import oracle
from models import chain, chain_states, golden, str_to_bytes
from quanty_amd import det as D
from quanty_amd import operators as O
from quanty_amd import problems as P
from quanty_amd import default_context
from quanty_amd._lib import QS_PARAM_ELL_PACK, QS_PARAM_SHIFT_MIN
from quanty_amd.engine import Basis, Hop, Op, apply, cf_spectrum, dot, expect

pytestmark = pytest.mark.gpu


def cvec(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def oracle_apply_dense(nf, terms, basis_words, x, restrictions, rows: Basis):
    """Oracle push-form O*x (restricted) scattered onto the row basis."""
    psi = oracle.Vec.from_arrays(nf, D.words_to_bytes(basis_words, nf), x)
    d, v = oracle.Op(nf, terms).apply(psi, restrictions=restrictions).arrays()
    out = np.zeros(rows.n, dtype=complex)
    if len(v):
        idx = rows.index(D.bytes_to_words(d, nf))
        assert np.all(idx >= 0)
        out[idx] = v
    return out


SMALL = [
    dict(seed=3, nbath=1, nelec=12, full_baths=1, max_holes=3, max_elec=0, d_range=(0, 10)),
    dict(seed=5, nbath=2, nelec=18, full_baths=1, max_holes=2, max_elec=2),
    dict(seed=9, nbath=3, nelec=22, full_baths=1, max_holes=2, max_elec=3),
]


@pytest.mark.parametrize("kw", SMALL)
def test_enumeration_bit_exact(kw):
    pr = P.anderson(**kw)
    nf = pr["NF"]
    b = Basis.enumerate(nf, pr["restrictions"])
    ref = oracle.enumerate_sector(nf, pr["restrictions"])
    assert b.n == len(ref)
    assert np.array_equal(b.words(), D.bytes_to_words(ref, nf))
    # rank lookup: every determinant maps to its position, foreign ones to -1
    assert np.array_equal(b.index(b.words()), np.arange(b.n))
    foreign = b.words().copy()
    foreign[:, -1] ^= np.uint64(1) << np.uint64(64 * b.nw - nf)  # flip the last mode
    idx = b.index(foreign)
    ok = np.array([all(lo <= bin(D.words_to_int(w) >> (64 * b.nw - nf) & int(m.replace(" ", ""), 2)).count("1") <= hi
                       for m, lo, hi in pr["restrictions"]) for w in foreign[:500]])
    assert np.array_equal(idx[:500] >= 0, ok)


def test_enumeration_uint128_and_overlapping_masks():
    nf = 76
    restr = [("1" * nf, 4, 4), ("1" * 20 + "0" * 56, 1, 3), ("0" * 10 + "1" * 60 + "0" * 6, 1, 3),
             ("0" * 70 + "1" * 6, 0, 1)]
    b = Basis.enumerate(nf, restr)
    assert b.nw == 2
    assert b.n == Basis.count(nf, restr) == 634745
    w = b.words()
    assert np.array_equal(w, D.bytes_to_words(oracle.enumerate_sector(nf, restr), nf))
    ints = [D.words_to_int(x) for x in w[:: max(1, b.n // 2000)]]
    assert all(x < y for x, y in zip(ints, ints[1:]))
    for x in ints:
        s = format(x >> (128 - nf), f"0{nf}b")
        for m, lo, hi in restr:
            assert lo <= sum(1 for p, q in zip(s, m) if p == q == "1") <= hi
    assert np.array_equal(b.index(w[:4096]), np.arange(min(4096, b.n)))


@pytest.mark.parametrize("kw", SMALL)
@pytest.mark.parametrize("mode", ["matfree", "assembled"])
def test_hpsi_matches_oracle(kw, mode):
    pr = P.anderson(**kw)
    nf = pr["NF"]
    b = Basis.enumerate(nf, pr["restrictions"])
    op = Op(nf, pr["H"])
    hop = Hop(op, b, mode=mode)
    assert hop.stats()["assembled"] == (mode == "assembled")
    x = cvec(b.n, 11)
    y = hop.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    want = oracle_apply_dense(nf, pr["H"], b.words(), x, pr["restrictions"], b)
    assert np.max(np.abs(y - want)) <= 1e-13 * np.max(np.abs(want))


class assembled_layout:
    """Assemble operators with the given shift-block threshold and ELL entry
    format inside the block; the defaults come back afterwards."""

    def __init__(self, shift_min, pack=1):
        self.v = {QS_PARAM_SHIFT_MIN: shift_min, QS_PARAM_ELL_PACK: pack}

    def __enter__(self):
        c = default_context()
        self.old = {k: c.get_param(k) for k in self.v}
        for k, v in self.v.items():
            c.set_param(k, v)

    def __exit__(self, *exc):
        c = default_context()
        for k, v in self.old.items():
            c.set_param(k, v)


@pytest.mark.parametrize("shift_min,pack", [(1, 1), (8, 1), (8, 0), (65, 1), (65, 0), (3, 0)])
def test_hpsi_layouts_match_oracle(shift_min, pack):
    """Every shift-block threshold and both ELL entry formats give the oracle's
    H*psi (square and rectangular bindings); the layout counts add up."""
    pr = P.anderson(**SMALL[2])
    nf = pr["NF"]
    b = Basis.enumerate(nf, pr["restrictions"])
    with assembled_layout(shift_min, pack):
        hop = Hop(Op(nf, pr["H"]), b, mode="assembled")
    st = hop.stats()
    assert st["shift_min"] == shift_min and st["packed"] == bool(pack)
    if shift_min == 65:
        assert st["nblocks"] == 0 and st["block_hits"] == 0
    else:
        assert st["nblocks"] > 0 and st["block_hits"] >= shift_min * st["nblocks"]
    assert st["ell_slots"] >= st["nnz_offdiag"] - st["block_hits"]
    x = cvec(b.n, 11)
    y = hop.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    want = oracle_apply_dense(nf, pr["H"], b.words(), x, pr["restrictions"], b)
    assert np.max(np.abs(y - want)) <= 1e-13 * np.max(np.abs(want))
    # rectangular: the Ni L-edge transition between two sectors
    ni = P.ni_xas()
    gs = Basis.enumerate(26, ni["gs_restrictions"])
    fs = Basis.enumerate(26, ni["fs_restrictions"])
    xg = cvec(gs.n, 2)
    with assembled_layout(shift_min, pack):
        th = Hop(Op(26, ni["T"][0]), fs, gs, mode="assembled")
    yt = th.apply(torch.from_numpy(xg).cuda()).cpu().numpy()
    wt = oracle_apply_dense(26, ni["T"][0], gs.words(), xg, ni["fs_restrictions"], fs)
    assert np.max(np.abs(yt - wt)) <= 1e-13 * np.max(np.abs(wt))


@pytest.mark.parametrize("pack", [1, 0])
def test_hpsi_uint128_matches_oracle(pack):
    """Two-word determinants (76 modes): H_i and H_f of the 76-mode L-edge model
    on small sectors, and the dipole between them, against the oracle; packed
    and per-field entries."""
    pr = P.large76(seed=2)
    nf = pr["NF"]
    full = [x for b in pr["baths"][:3] for x in b]
    empty = [x for b in pr["baths"][3:] for x in b]
    common = [(P.mask(nf, full), 30, 30), (P.mask(nf, empty), 0, 1), (P.mask(nf, range(nf)), 44, 44)]
    gsr = [(P.mask(nf, pr["pmodes"]), 6, 6), (P.mask(nf, pr["dmodes"]), 7, 8)] + common
    fsr = [(P.mask(nf, pr["pmodes"]), 5, 5), (P.mask(nf, pr["dmodes"]), 8, 9)] + common
    gs, fs = Basis.enumerate(nf, gsr), Basis.enumerate(nf, fsr)
    assert gs.nw == 2 and (gs.n, fs.n) == (3645, 8160)
    with assembled_layout(8, pack):
        cases = [(pr["Hi"], gs, gs, gsr), (pr["Hf"], fs, fs, fsr), (pr["T"], fs, gs, fsr)]
        hops = [(Hop(Op(nf, t), r, c, mode=mode), t, r, c, rr) for t, r, c, rr in cases
                for mode in ("assembled", "matfree")]
    for hop, terms, rows, cols, rr in hops:
        if hop.stats()["assembled"]:
            assert hop.stats()["packed"] == bool(pack)
        x = cvec(cols.n, 5)
        y = hop.apply(torch.from_numpy(x).cuda()).cpu().numpy()
        want = oracle_apply_dense(nf, terms, cols.words(), x, rr, rows)
        assert np.max(np.abs(y - want)) <= 1e-13 * np.max(np.abs(want))


def test_optimespsi_known_answer_gpu():
    terms = [O.written([O.cr(0), O.an(0)], 1j), O.written([O.cr(0), O.cr(1), O.an(0), O.an(1)], 1.0)]
    b = Basis.enumerate(2, [("11", 2, 2)])
    y = apply(Op(2, terms), b, b, torch.tensor([1.0 + 0j], dtype=torch.complex128).cuda()).cpu().numpy()
    assert y[0] == complex(-1, 1)


def test_transition_between_sectors():
    """T maps the 2p6 ground sector to the 2p5 final sector (rows != cols)."""
    ni = P.ni_xas()
    gs = Basis.enumerate(26, ni["gs_restrictions"])
    fs = Basis.enumerate(26, ni["fs_restrictions"])
    x = cvec(gs.n, 2)
    for T in ni["T"]:
        for mode in ["matfree", "assembled"]:
            hop = Hop(Op(26, T), fs, gs, mode=mode)
            y = hop.apply(torch.from_numpy(x).cuda()).cpu().numpy()
            want = oracle_apply_dense(26, T, gs.words(), x, ni["fs_restrictions"], fs)
            assert np.max(np.abs(y - want)) <= 1e-13 * np.max(np.abs(want))


def test_hermiticity_and_expectation():
    pr = P.anderson(**SMALL[2])
    nf = pr["NF"]
    b = Basis.enumerate(nf, pr["restrictions"])
    op = Op(nf, pr["H"])
    hop = Hop(op, b, mode="assembled")
    x = torch.from_numpy(cvec(b.n, 1)).cuda()
    z = torch.from_numpy(cvec(b.n, 2)).cuda()
    lhs = dot(z, hop.apply(x))
    rhs = dot(hop.apply(z), x)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    e = expect(op, b, x)
    assert abs(e - dot(x, hop.apply(x))) < 1e-12 * abs(e)
    assert abs(e.imag) < 1e-10 * abs(e)


@pytest.mark.parametrize("kw", SMALL[1:])
@pytest.mark.parametrize("mode", ["matfree", "assembled"])
def test_lanczos_matches_oracle(kw, mode):
    pr = P.anderson(**kw)
    nf = pr["NF"]
    b = Basis.enumerate(nf, pr["restrictions"])
    hop = Hop(Op(nf, pr["H"]), b, mode=mode)
    x = cvec(b.n, 4)
    a, bb = hop.lanczos(torch.from_numpy(x).cuda(), 40)
    psi = oracle.Vec.from_arrays(nf, D.words_to_bytes(b.words(), nf), x)
    ra, rb = oracle.Op(nf, pr["H"]).lanczos(psi, 40, restrictions=pr["restrictions"])
    assert np.allclose(a, ra, rtol=1e-10, atol=1e-12)
    assert np.allclose(bb, rb, rtol=1e-10, atol=0)


def _many_term_operator(nf, nterms, seed):
    """Hermitian two-body operator with more distinct terms than the LDS
    staging of the term values takes, so the assembled kernels read them from
    global memory."""
    rng = np.random.default_rng(seed)
    terms = []
    seen = set()
    while len(terms) < nterms:
        i, j, k, l = (int(q) for q in rng.integers(0, nf, 4))
        if i == j or k == l or (i, j, k, l) in seen:
            continue
        seen.add((i, j, k, l))
        terms.append(O.written([O.cr(i), O.cr(j), O.an(k), O.an(l)], complex(*rng.normal(size=2))))
    return O.add(terms, O.dagger(terms))


def test_many_terms_global_coef_path():
    """H*psi, Lanczos and the interleaved block product when the term table is
    too large for LDS (2000+ terms on a 16-mode, 8-electron basis)."""
    nf = 16
    terms = _many_term_operator(nf, 1100, 21)
    restr = [("1" * nf, 8, 8)]
    b = Basis.enumerate(nf, restr)
    op = Op(nf, terms)
    assert op.nterms > 1024
    hop = Hop(op, b, mode="assembled")
    x = cvec(b.n, 3)
    y = hop.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    want = oracle_apply_dense(nf, terms, b.words(), x, restr, b)
    assert np.max(np.abs(y - want)) <= 1e-13 * np.max(np.abs(want))
    a, bb = hop.lanczos(torch.from_numpy(x).cuda(), 30)
    psi = oracle.Vec.from_arrays(nf, D.words_to_bytes(b.words(), nf), x)
    ra, rb = oracle.Op(nf, terms).lanczos(psi, 30, restrictions=restr)
    assert np.allclose(a, ra, rtol=1e-10, atol=1e-12)
    assert np.allclose(bb, rb, rtol=1e-10, atol=0)
    # 3 orthonormal start vectors (the first Krylov block is then the start
    # block itself): A_1 = V^+ H V through the 3-vector kernel
    q, _ = np.linalg.qr(np.stack([cvec(b.n, 30 + v) for v in range(3)], axis=1))
    A, _, kept = hop.block_lanczos(torch.from_numpy(np.ascontiguousarray(q.T)).cuda(), 1)
    assert kept[1] == 3
    hq = np.stack([oracle_apply_dense(nf, terms, b.words(), q[:, v], restr, b) for v in range(3)], axis=1)
    A1 = q.conj().T @ hq
    assert np.max(np.abs(A[1] - A1)) <= 1e-11 * np.max(np.abs(A1))


@pytest.mark.parametrize("name", ["G0Dn1", "G0Dn2", "G0Up1", "G0Up2", "G1Dn1", "G1Dn2", "G1Up1", "G1Up2",
                                  "G2Dn1", "G2Dn2", "G2Up1", "G2Up2"])
def test_lanczos_golden_tri_tables_gpu(name):
    """GPU Lanczos on the restricted 6-site chain (TestScripts/CreateSpectra.Quanty:596-760):
    golden a/b and the G(w, 0.1) checks at 1e-13."""
    g = golden("createspectra_tri.json")
    ref = g["tables"][name]
    nf, terms = chain(6, int(name[1]))
    restr = [("00" + "0" * 10 + "11", 0, 0)]
    strings, coefs = chain_states(nf, int(name[-1]))
    gs = Basis.enumerate(nf, restr + [("1" * nf, 1, 1)])
    fs = Basis.enumerate(nf, restr + [("1" * nf, 2, 2)])
    x = np.zeros(gs.n, dtype=complex)
    idx = gs.index(np.array([D.string_to_words(s, nf) for s in strings], dtype=np.uint64))
    x[idx] = coefs
    T = [((O.cr(0 if name[2:4] == "Dn" else 1),), 1.0)]
    tx = Hop(Op(nf, T), fs, gs, mode="matfree").apply(torch.from_numpy(x).cuda())
    hop = Hop(Op(nf, terms), fs, mode="assembled")
    a, b = hop.lanczos(tx, 200)
    na, nb = len(ref["a"]), len(ref["b"])
    assert np.allclose(a[:na], ref["a"], atol=5e-13, rtol=1e-12)
    assert np.allclose(b[:nb], ref["b"], atol=5e-13, rtol=1e-12)
    for w in g["omega"]:
        got = cf_spectrum(a, b, 0.0, w, w + 1.0, 1, g["gamma"])[0, 0]
        want = oracle.cf(np.array(ref["a"]), np.array(ref["b"]), 0.0, w, w + 1.0, 1, g["gamma"])[0]
        assert abs(got - want) < g["threshold_block2"]


def test_cf_matches_oracle():
    rng = np.random.default_rng(5)
    nspec, n, ne = 3, 150, 2000
    a = rng.normal(size=(nspec, n + 1))
    b = np.abs(rng.normal(size=(nspec, n))) + 0.05
    E0 = rng.normal(size=nspec)
    g = cf_spectrum(a, b, E0, -10.0, 20.0, ne, 0.4)
    for s in range(nspec):
        r = oracle.cf(a[s], b[s], E0[s], -10.0, 20.0, ne, 0.4)
        assert np.max(np.abs(g[s] - r)) <= 1e-13 * np.max(np.abs(r))


def test_anderson_full_size_properties():
    """The 50-mode, 2,179,235-determinant workload: assembled == matrix-free,
    Hermiticity, and columns H e_j equal to the oracle's push-form columns."""
    pr = P.anderson(seed=1)
    nf = pr["NF"]
    b = Basis.enumerate(nf, pr["restrictions"])
    assert b.n == 2179235
    op = Op(nf, pr["H"])
    asm = Hop(op, b, mode="assembled")
    mf = Hop(op, b, mode="matfree")
    x = torch.from_numpy(cvec(b.n, 8)).cuda()
    y1 = asm.apply(x)
    y2 = mf.apply(x)
    assert (torch.max(torch.abs(y1 - y2)) / torch.max(torch.abs(y2))).item() < 1e-13
    z = torch.from_numpy(cvec(b.n, 9)).cuda()
    lhs, rhs = dot(z, y1), dot(asm.apply(z), x)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    words = b.words()
    rng = np.random.default_rng(0)
    for j in rng.integers(0, b.n, size=6):
        e = torch.zeros(b.n, dtype=torch.complex128, device="cuda")
        e[int(j)] = 1.0
        col = asm.apply(e).cpu().numpy()
        want = oracle_apply_dense(nf, pr["H"], words[j:j + 1], np.array([1.0 + 0j]), pr["restrictions"], b)
        assert np.max(np.abs(col - want)) < 1e-13 * np.max(np.abs(want))
    st = asm.stats()
    assert st["nnz_eff"] > 0 and st["nnz_stored"] >= 0


def test_hpsi_complex_diagonal_matches_oracle():
    """A square operator whose diagonal terms are complex (not Hermitian) keeps
    its diagonal as complex numbers; the Hermitian H beside it stores a real
    one. Both against the oracle, assembled and matrix-free."""
    pr = P.anderson(**SMALL[1])
    nf = pr["NF"]
    b = Basis.enumerate(nf, pr["restrictions"])
    rng = np.random.default_rng(5)
    # number operators n_i with complex weights, plus H's off-diagonal part
    nterms = [((0x8000 | m, m), complex(rng.normal(), rng.normal())) for m in range(nf)]
    terms = list(pr["H"]) + nterms
    x = cvec(b.n, 17)
    stream = {}
    for tl, complex_diag in ((pr["H"], False), (terms, True)):
        want = oracle_apply_dense(nf, tl, b.words(), x, pr["restrictions"], b)
        for mode in ("assembled", "matfree"):
            hop = Hop(Op(nf, tl), b, mode=mode)
            y = hop.apply(torch.from_numpy(x).cuda()).cpu().numpy()
            assert np.max(np.abs(y - want)) <= 1e-13 * np.max(np.abs(want)), (mode, complex_diag)
            if mode == "assembled":
                stream[complex_diag] = hop.stats()["stream_bytes"]
    # the added terms are diagonal: the off-diagonal stream is the same, the
    # diagonal takes 16 instead of 8 bytes per row
    assert stream[True] - stream[False] == 8 * b.n
