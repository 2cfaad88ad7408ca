"""The oracle (CPU restatement) pinned against the reference's golden data and
against mathematics that fixes the answer independently of the restatement."""
import math

import numpy as np
import pytest

import oracle
from models import chain, chain_states, dense_from_terms, f_shell, golden, jw_matrices, str_to_bytes
from quanty_amd import operators as O


def vec_from_strings(nf, strings, coefs):
    v = oracle.Vec(nf)
    for s, c in zip(strings, coefs):
        v.add(str_to_bytes(s, nf), c)
    return v


# ---------------------------------------------------------------------------
# signs: known-answer test and Jordan-Wigner brute force
# ---------------------------------------------------------------------------
def test_optimespsi_known_answer():
    """Debug examples/Reference/Objects/Operator/OpTimesPsi.Quanty:
    (i n0 + c+0 c+1 c0 c1)|11> = (-1+i)|11>."""
    terms = [O.written([O.cr(0), O.an(0)], 1j), O.written([O.cr(0), O.cr(1), O.an(0), O.an(1)], 1.0)]
    op = oracle.Op(2, terms)
    out = op.apply(vec_from_strings(2, ["11"], [1.0]))
    d, v = out.arrays()
    assert out.size() == 1
    assert d[0, 0] == 0b11000000
    assert v[0] == complex(-1, 1)


@pytest.mark.parametrize("nf", [5, 8, 9])
def test_ladder_matches_jordan_wigner(nf):
    """Every ladder string of length <= 4 on every determinant: phase and target
    equal the Jordan-Wigner matrix element."""
    rng = np.random.default_rng(nf)
    cs = jw_matrices(nf)
    dim = 2 ** nf
    for _ in range(60):
        L = int(rng.integers(1, 5))
        codes = [int(rng.integers(0, nf)) | (0x8000 if rng.integers(0, 2) else 0) for _ in range(L)]
        M = np.eye(dim)
        for c in codes:
            M = (cs[c & 0x7FFF].T if c & 0x8000 else cs[c & 0x7FFF]) @ M
        for idx in rng.integers(0, dim, size=24):
            bits = format(int(idx), f"0{nf}b")
            ph, out = oracle.ladder(str_to_bytes(bits, nf), codes)
            col = M[:, idx]
            nz = np.nonzero(col)[0]
            if len(nz) == 0:
                assert ph == 0.0
            else:
                assert ph == col[nz[0]]
                tgt = format(int(nz[0]), f"0{nf}b")
                assert np.array_equal(out, str_to_bytes(tgt, nf))


def test_opsi_matches_dense_jordan_wigner():
    """O*psi of a random complex operator with one- and two-body terms equals the
    dense Jordan-Wigner product, and the restricted variant equals the dense
    product projected on the restricted set."""
    nf = 8
    rng = np.random.default_rng(3)
    terms = []
    for _ in range(40):
        L = int(rng.choice([0, 2, 4]))
        codes = [int(rng.integers(0, nf)) | (0x8000 if k >= L // 2 else 0) for k in range(L)]
        terms.append((tuple(codes), complex(rng.normal(), rng.normal())))
    H = dense_from_terms(nf, terms)
    psi_full = np.zeros(2 ** nf, dtype=complex)
    idx = rng.choice(2 ** nf, 30, replace=False)
    psi_full[idx] = rng.normal(size=30) + 1j * rng.normal(size=30)
    psi = oracle.Vec(nf)
    for i in idx:
        psi.add(str_to_bytes(format(int(i), f"0{nf}b"), nf), psi_full[i])
    want = H @ psi_full
    out = oracle.Op(nf, terms).apply(psi)
    d, v = out.arrays()
    got = np.zeros_like(want)
    for k in range(d.shape[0]):
        got[int(d[k, 0])] += v[k]
    assert np.allclose(got, want, atol=1e-12, rtol=0)
    # restricted: modes 0..3 hold exactly 2 particles
    restr = [("11110000", 2, 2)]
    out = oracle.Op(nf, terms).apply(psi, restrictions=restr)
    d, v = out.arrays()
    got = np.zeros_like(want)
    for k in range(d.shape[0]):
        got[int(d[k, 0])] += v[k]
    keep = np.array([bin(i >> 4).count("1") == 2 for i in range(2 ** nf)])
    assert np.allclose(got, np.where(keep, want, 0), atol=1e-12, rtol=0)


def test_restriction_filter_counts_bits():
    nf = 12
    m = "110011001100"
    for idx in range(0, 2 ** nf, 37):
        s = format(idx, f"0{nf}b")
        n = sum(1 for a, b in zip(s, m) if a == b == "1")
        masks = np.zeros((1, 2), dtype=np.uint8)
        masks[0] = str_to_bytes(m, nf)
        det = str_to_bytes(s, nf)
        lo = np.array([2], dtype=np.int32)
        hi = np.array([3], dtype=np.int32)
        ok = oracle.load().or_restricted_ok(det.ctypes.data, 2, 1, masks.ctypes.data, lo.ctypes.data,
                                            hi.ctypes.data)
        assert ok == (2 <= n <= 3)


# ---------------------------------------------------------------------------
# sectors
# ---------------------------------------------------------------------------
def test_enumeration_brute_force():
    nf = 14
    restr = [("11111111111111", 7, 7), ("11110000001111", 2, 5), ("00111111110000", 3, 8),
             ("10101010101010", 1, 4)]
    got = oracle.enumerate_sector(nf, restr)
    want = []
    for idx in range(2 ** nf):
        s = format(idx, f"0{nf}b")
        if all(lo <= sum(1 for a, b in zip(s, m) if a == b == "1") <= hi for m, lo, hi in restr):
            want.append(str_to_bytes(s, nf))
    assert np.array_equal(got, np.array(want))


def test_sector_sizes_closed_form():
    """Sizes fixed by combinatorics: sum over group occupations of products of binomials."""
    from quanty_amd import problems as P
    ni = P.ni_xas()
    assert len(oracle.enumerate_sector(26, ni["gs_restrictions"])) == math.comb(20, 18) == 190
    assert len(oracle.enumerate_sector(26, ni["fs_restrictions"])) == 6 * math.comb(20, 19) == 120
    # Anderson impurity, 50 modes: d in [6,10], B1+B2 holes <= 2, B3+B4 electrons <= 2, N = 28
    want = 0
    for nd in range(6, 11):
        for h in range(0, 3):
            e = 28 - nd - (20 - h)
            if 0 <= e <= 2:
                want += math.comb(10, nd) * math.comb(20, h) * math.comb(20, e)
    assert want == 2179235
    from quanty_amd.engine import Basis
    assert Basis.count(50, P.anderson(seed=1)["restrictions"]) == want


# ---------------------------------------------------------------------------
# Lanczos and the continued fraction
# ---------------------------------------------------------------------------
def test_cf_equals_resolvent():
    """The continued fraction of (a, b) equals a0 + b0^2 <e1|(z - T)^-1|e1>."""
    rng = np.random.default_rng(0)
    n = 12
    a = np.concatenate([[0.3], rng.normal(size=n)])
    b = np.abs(rng.normal(size=n)) + 0.1
    T = np.diag(a[1:]) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
    E0, emin, emax, ne, gamma = 0.7, -4.0, 4.0, 40, 0.3
    g = oracle.cf(a, b, E0, emin, emax, ne, gamma)
    w = emin + np.arange(ne + 1) * (emax - emin) / ne
    for i in range(ne + 1):
        z = w[i] + E0 + 0.5j * gamma
        r = np.linalg.solve(z * np.eye(n) - T, np.eye(n)[:, 0])[0]
        assert abs(g[i] - (a[0] + b[0] ** 2 * r)) < 1e-12


def test_lanczos_exhausting_krylov_gives_exact_resolvent():
    """With as many steps as the Krylov space has dimensions, G(z) from the
    Lanczos coefficients equals <v|(z-H)^-1|v> from the dense Hamiltonian."""
    nf, terms = chain(nbath=3, kind=1)
    H = dense_from_terms(nf, terms)
    strings, coefs = chain_states(nf, 2)
    # T = c+_0, psi2 -> T psi lives in the 2-particle sector
    psi = np.zeros(2 ** nf, dtype=complex)
    for s, c in zip(strings, coefs):
        psi[int(s, 2)] = c
    cs = jw_matrices(nf)
    tpsi = cs[0].T @ psi
    v = oracle.Vec(nf)
    for i in np.nonzero(tpsi)[0]:
        v.add(str_to_bytes(format(int(i), f"0{nf}b"), nf), tpsi[i])
    sector = [i for i in range(2 ** nf) if bin(i).count("1") == 2]
    a, b = oracle.Op(nf, terms).lanczos(v, len(sector))
    for z in [-1.1 + 0.05j, 0.3 + 0.05j, 2.5 + 0.05j]:
        r = tpsi.conj() @ np.linalg.solve(z * np.eye(2 ** nf) - H, tpsi)
        g = oracle.cf(a, b, 0.0, z.real, z.real + 1.0, 1, 2 * z.imag)[0]
        assert abs(g - r) < 1e-9 * max(1.0, abs(r))


@pytest.mark.parametrize("name", ["G0Dn1", "G0Dn2", "G0Up1", "G0Up2", "G1Dn1", "G1Dn2", "G1Up1", "G1Up2",
                                  "G2Dn1", "G2Dn2", "G2Up1", "G2Up2"])
@pytest.mark.parametrize("block", [1, 2])
def test_lanczos_golden_tri_tables(name, block):
    """The oracle reproduces the reference's golden Tri tables
    (TestScripts/CreateSpectra.Quanty:188-260, 664-736) and the G(w, 0.1)
    checks made there, using the reference defaults NTri=200 and
    epsilon = sqrt(DBL_EPSILON)/100, Gamma = max(0, epsilon)."""
    g = golden("createspectra_tri.json")
    ref = g["tables"][name]
    kind = int(name[1])
    op_mode = 0 if name[2:4] == "Dn" else 1
    which = int(name[-1])
    nbath = 5 if block == 1 else 6
    restr = None if block == 1 else [("00" + "0" * 10 + "11", 0, 0)]
    nf, terms = chain(nbath, kind)
    strings, coefs = chain_states(nf, which)
    psi = vec_from_strings(nf, strings, coefs)
    tpsi = oracle.Op(nf, [((O.cr(op_mode),), 1.0)]).apply(psi, restrictions=restr)
    eps = math.sqrt(np.finfo(float).eps) / 100
    a, b = oracle.Op(nf, terms).lanczos(tpsi, 200, epsilon=eps, gamma=eps, restrictions=restr)
    na, nb = len(ref["a"]), len(ref["b"])
    assert np.allclose(a[:na], ref["a"], atol=5e-13, rtol=1e-12)
    assert np.allclose(b[:nb], ref["b"], atol=5e-13, rtol=1e-12)
    thr = g["threshold_block1"] if block == 1 else g["threshold_block2"]
    for w in g["omega"]:
        got = oracle.cf(a, b, 0.0, w, w + 1.0, 1, g["gamma"])[0]
        want = oracle.cf(np.array(ref["a"]), np.array(ref["b"]), 0.0, w, w + 1.0, 1, g["gamma"])[0]
        assert abs(got - want) < thr


@pytest.mark.parametrize("cplx", [False, True])
def test_f6_mathematica_energies(cplx):
    """Dense diagonalization of the oracle's H on the 3003-dim f6 sector gives the
    Mathematica energies of TestScripts/Eigensystem.Quanty:571/669 (pins the
    Coulomb/ldots construction and the oracle's H*psi)."""
    nf, terms = f_shell(cplx)
    sector = oracle.enumerate_sector(nf, [("1" * nf, 6, 6)])
    n = len(sector)
    assert n == 3003
    op = oracle.Op(nf, terms)
    key = {bytes(sector[i]): i for i in range(n)}
    H = np.zeros((n, n), dtype=complex)
    for j in range(n):
        v = oracle.Vec(nf)
        v.add(sector[j], 1.0)
        d, val = op.apply(v).arrays()
        for k in range(d.shape[0]):
            H[key[bytes(d[k])], j] += val[k]
    assert np.allclose(H, H.conj().T, atol=1e-13)
    e = np.linalg.eigvalsh(H)[:10]
    want = np.array(golden("eigensystem_f6.json")["complex" if cplx else "real"])
    if not cplx:
        assert np.allclose(e, want, rtol=1e-10, atol=0)
    else:
        # the reference holds its complex-case values only to 1e-4 per level and
        # 1e-6 for the ground state (TestScripts/Eigensystem.Quanty:684,691)
        assert np.all(np.abs(e - want) < 1e-4)
        assert abs(e[0] - want[0]) < 1e-6


# ---------------------------------------------------------------------------
# broadening: or_broaden / or_fwhm_table
# ---------------------------------------------------------------------------
def _delta(npts, i0):
    G = np.zeros(npts, dtype=complex)
    G[i0] = 1.0 - 0.5j
    return G


def test_gaussian_broadening_conserves_weight_and_has_the_fwhm_variance():
    """A Gaussian of FWHM f has sigma = f / (2 sqrt(2 ln 2)); averaging over the 7
    sub-points j/7 of the cell adds dE^2 * mean(j^2)/49 = 4 dE^2 / 49 to the
    variance. The Riemann sums of a well-resolved Gaussian are exact to
    exp(-2 pi^2 sigma^2 / dE^2), so weight and moments are pinned to rounding."""
    npts, dE, f = 4001, 0.01, 0.5
    i0 = 2000
    out = oracle.broaden(_delta(npts, i0), dE, "G", f)[0]
    m = (np.arange(npts) - i0) * dE
    sigma = f / (2 * math.sqrt(2 * math.log(2)))
    w = out / (1.0 - 0.5j)
    assert np.max(np.abs(w.imag)) < 1e-15
    w = w.real
    assert abs(w.sum() - 1) < 1e-12
    assert abs((m * w).sum()) < 1e-14
    assert abs((m * m * w).sum() - (sigma ** 2 + 4 * dE ** 2 / 49)) < 1e-12 * sigma ** 2
    # a narrower FWHM than half the step leaves the spectrum unchanged
    G = np.random.default_rng(1).standard_normal((1, 50)) + 0j
    assert np.array_equal(oracle.broaden(G, 0.1, "G", 0.049), G)


def test_lorentzian_broadening_adds_widths():
    """Lorentzian (FWHM g2) convolved with the line of a pole broadened by g1 is
    the pole's line with g1 + g2 (closed form; Re and Im alike). The grid
    discretization and the cut tails leave a few 1e-4 near the pole."""
    npts, emin, emax = 6001, -30.0, 30.0
    dE = (emax - emin) / (npts - 1)
    g1, g2, E1 = 0.3, 0.4, 0.7
    a, b = np.array([0.0, E1]), np.array([1.0])
    G1 = oracle.cf(a, b, 0.0, emin, emax, npts - 1, g1)
    G12 = oracle.cf(a, b, 0.0, emin, emax, npts - 1, g1 + g2)
    out = oracle.broaden(G1, dE, "L", g2)[0]
    x = emin + np.arange(npts) * dE
    near = np.abs(x - E1) < 3
    scale = np.max(np.abs(G12))
    assert np.max(np.abs(out[near] - G12[near])) < 3e-4 * scale
    # doubling or halving the width would be far off
    wrong = oracle.broaden(G1, dE, "L", 2 * g2)[0]
    assert np.max(np.abs(wrong[near] - G12[near])) > 0.05 * scale


@pytest.mark.parametrize("kind", ["G", "L"])
def test_variable_width_uses_the_source_point_width(kind):
    """Per-point widths: a delta at i0 spreads with the width of i0, whatever the
    widths elsewhere; constant per-point widths equal the single width
    bit for bit; widths below dE/2 are raised to dE/2."""
    npts, dE, i0 = 801, 0.02, 300
    rng = np.random.default_rng(3)
    f = rng.uniform(0.1, 1.0, npts)
    out = oracle.broaden(_delta(npts, i0), dE, kind, f, per_point=True)
    ref = oracle.broaden(_delta(npts, i0), dE, kind, f[i0])
    assert np.array_equal(out, ref)
    G = rng.standard_normal((2, npts)) + 1j * rng.standard_normal((2, npts))
    assert np.array_equal(oracle.broaden(G, dE, kind, np.full(npts, 0.3), per_point=True),
                          oracle.broaden(G, dE, kind, 0.3))
    assert np.array_equal(oracle.broaden(G, dE, kind, np.full(npts, 0.001), per_point=True),
                          oracle.broaden(G, dE, kind, dE / 2))


def test_fwhm_table_is_linear_interpolation():
    rng = np.random.default_rng(5)
    emin, npts = -3.0, 601
    dE = 0.01
    for nt in (1, 2, 5):
        E = np.sort(rng.uniform(-4, 4, nt))
        E[0] = emin + 17 * dE  # a table node on a grid point
        E.sort()
        F = rng.uniform(0.1, 1, nt)
        got = oracle.fwhm_table(np.stack([E, F], 1), emin, dE, npts)
        x = emin + np.arange(npts) * dE
        assert np.allclose(got, np.interp(x, E, F), rtol=0, atol=1e-14)


def test_hubbard_ring_energies_golden():
    """The 10 lowest energies of the 8-site Hubbard ring held by
    TestScripts/Eigensystem.Quanty:1119, from the oracle's push-form H*psi in a
    matrix-free Lanczos (scipy eigsh) on the 4 down + 4 up sector (4900 dets).
    The golden values carry Quanty's own convergence error (two degenerate pairs
    differ by 2e-10 and 1e-8), so they are compared at 1e-7."""
    from scipy.sparse.linalg import LinearOperator, eigsh

    from models import hubbard_ring
    g = golden("eigensystem_hubbard.json")
    nf, terms = hubbard_ring(g["nsites"], g["t1"], g["t2"], g["U"])
    restr = [("1010101010101010", 4, 4), ("0101010101010101", 4, 4)]
    dets = oracle.enumerate_sector(nf, restr)
    assert len(dets) == 4900
    index = {bytes(d): i for i, d in enumerate(dets)}
    op = oracle.Op(nf, terms)

    def mv(x):
        x = np.asarray(x).ravel()
        d, v = op.apply(oracle.Vec.from_arrays(nf, dets, x.astype(complex))).arrays()
        y = np.zeros(len(dets), dtype=complex)
        for det, val in zip(d, v):
            y[index[bytes(det)]] += val
        return y.real if np.isrealobj(x) else y

    H = LinearOperator((4900, 4900), matvec=mv, dtype=float)
    ev = np.sort(eigsh(H, k=10, which="SA", ncv=48, tol=1e-13, return_eigenvectors=False))
    assert np.max(np.abs(ev - np.array(g["energies"]))) < 1e-7


def _sector_ops(nf, dets, restr):
    """(apply, to_vec) over a sorted-by-key determinant list: to_vec places the
    oracle's output determinants on the list's positions."""
    def key(d):
        d = np.asarray(d, dtype=np.uint64)
        k = np.zeros(len(d), dtype=np.uint64)
        for j in range(d.shape[1]):
            k = (k << np.uint64(8)) | d[:, j]
        return k

    kd = key(dets)
    order = np.argsort(kd)
    ks = kd[order]
    n = len(dets)

    def to_vec(d, v):
        k = key(d)
        pos = np.minimum(np.searchsorted(ks, k), n - 1)
        assert np.all(ks[pos] == k)
        y = np.zeros(n, dtype=complex)
        np.add.at(y, order[pos], v)
        return y

    def apply(op, x, restricted=False):
        psi = oracle.Vec.from_arrays(nf, dets, np.asarray(x, dtype=complex))
        return op.apply(psi, restrictions=restr if restricted else None).arrays()

    return apply, to_vec


@pytest.mark.parametrize("axis", ["z", "y"])
def test_mlft_d5_expectation_values_golden(axis):
    """The d5 ligand-field cluster of TestScripts/Eigensystem.Quanty:750-1012: the
    6 lowest states (a Zeeman-split 6A1 sextet, 2e-6 apart) from the oracle's
    restricted H*psi in scipy eigsh, and their 14 expectation values (energy, the
    cluster's S^2, L^2, J^2 as sums of squared component norms, 3d spin, orbital
    and spin-orbit moments, F2 and F4 energies, eg/t2g occupations of both shells,
    3d count) against the table of :763-770, which the script checks to 2e-6;
    with the field along y (a complex Hamiltonian) against :1032-1039.
    Pins the shell operators, Coulomb, cubic crystal field and hybridization
    construction as well as the restricted push H*psi."""
    from scipy.sparse.linalg import LinearOperator, eigsh

    from models import mlft_d5
    g = golden("eigensystem_mlft.json")
    nf, H, obs, comps = mlft_d5(axis)
    want = g["values"] if axis == "z" else g["values_y"]
    names = ["S%s_3d" % axis, "L%s_3d" % axis] + g["observables"][6:]
    restr = [("1" * 20, 15, 15), ("1" * 10 + "0" * 10, 5, 7)]
    dets = oracle.enumerate_sector(nf, restr)
    assert len(dets) == 252 + 2100 + 5400
    apply, to_vec = _sector_ops(nf, dets, restr)
    h = oracle.Op(nf, H)

    dtype = float if axis == "z" else complex

    def mv(x):
        y = to_vec(*apply(h, np.asarray(x).ravel(), restricted=True))
        return y.real if dtype is float else y

    w, V = eigsh(LinearOperator((len(dets),) * 2, matvec=mv, dtype=dtype), k=6, which="SA", ncv=40, tol=1e-14)
    ops = {k: oracle.Op(nf, t) for k, t in obs.items()}
    cops = {k: [oracle.Op(nf, c) for c in v] for k, v in comps.items()}
    for r, i in enumerate(np.argsort(w)):
        x = V[:, i].astype(complex)
        row = [w[i]]
        for k in "SLJ":
            row.append(sum(np.vdot(v, v).real for _, v in (apply(c, x) for c in cops[k])))
        for k in names:
            row.append(np.vdot(x, to_vec(*apply(ops[k], x))).real)
        assert np.max(np.abs(np.array(row) - np.array(want[r]))) < 2e-8, r


def test_threaded_lanczos_keeps_orthonormal_vectors():
    """or_lanczos_mt: with several threads the coefficients equal the
    single-thread recurrence to rounding, the kept vectors are the normalized
    Lanczos vectors (orthonormal while the recurrence keeps orthogonality,
    v_0 = psi/|psi|), and or_combine forms sum z_i v_i: the Ritz vector of the
    lowest eigenvalue of the table has the Rayleigh quotient that eigenvalue."""
    import oracle
    from models import chain
    nf, terms = chain(5, 1)
    op = oracle.Op(nf, terms)
    rng = np.random.default_rng(2)
    dets = oracle.enumerate_sector(nf, [("1" * nf, 6, 6)])
    psi = oracle.Vec.from_arrays(nf, dets, rng.normal(size=len(dets)) + 1j * rng.normal(size=len(dets)))
    a1, b1 = op.lanczos(psi, 12)
    a2, b2, keep = op.lanczos(psi, 12, nthreads=4, keep=True)
    assert np.allclose(a1, a2, rtol=1e-12, atol=1e-12) and np.allclose(b1, b2, rtol=1e-12)
    G = np.array([[x.dot(y) for y in keep] for x in keep])
    assert np.max(np.abs(G - np.eye(len(keep)))) < 1e-10
    d0, c0 = psi.arrays()
    _, k0 = keep[0].arrays()
    assert np.allclose(k0, c0 / np.linalg.norm(c0), atol=1e-14)
    T = np.diag(a2[1:]) + np.diag(b2[1:], 1) + np.diag(b2[1:], -1)
    w, z = np.linalg.eigh(T)
    y = oracle.combine(keep, z[:, 0])
    hy = op.apply(y)
    assert abs(y.dot(hy).real / y.dot(y).real - w[0]) < 1e-9
