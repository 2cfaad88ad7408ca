"""Front-end semantics that need no GPU: operator parsing and algebra, options."""
import numpy as np
import pytest

from models import dense_from_terms
from quanty_amd import QSError
from quanty_amd.api import NewOperator, NewWavefunction, Operator


def test_raw_operator_order_and_sign():
    """{i, -j, v} is v c+_i c_j: the rightmost written factor acts first; -0.0 annihilates mode 0."""
    O = NewOperator(3, 0, [[2, -0.0, 0.5], [0, -0.0, 1]])
    assert O.terms[(0, 0x8000 | 2)] == 0.5
    assert O.terms[(0, 0x8000 | 0)] == 1
    H = dense_from_terms(3, list(O.terms.items()))
    # c+_2 c_0 |100> = |001> with sign (+1): mode 0 is bit 2 (most significant) of the index
    assert H[0b001, 0b100] == 0.5


def test_operator_algebra_matches_dense():
    nf = 6
    rng = np.random.default_rng(2)
    A = NewOperator(nf, 0, [[int(rng.integers(0, nf)), -float(rng.integers(0, nf)), complex(*rng.normal(size=2))]
                            for _ in range(8)])
    B = NewOperator(nf, 0, [[int(rng.integers(0, nf)), int(rng.integers(0, nf)), -float(rng.integers(0, nf)),
                             -float(rng.integers(0, nf)), complex(*rng.normal(size=2))] for _ in range(8)])
    dA, dB = dense_from_terms(nf, list(A.terms.items())), dense_from_terms(nf, list(B.terms.items()))
    assert np.allclose(dense_from_terms(nf, list((A * B).terms.items())), dA @ dB)
    assert np.allclose(dense_from_terms(nf, list((A + 2j * B).terms.items())), dA + 2j * dB)
    assert np.allclose(dense_from_terms(nf, list(A.ConjugateTranspose().terms.items())), dA.conj().T)
    assert np.allclose(dense_from_terms(nf, list((A - 1.5).terms.items())), dA - 1.5 * np.eye(2 ** nf))


def test_named_operators():
    nf = 10
    up, dn = [1, 3, 5, 7, 9], [0, 2, 4, 6, 8]
    Lz = dense_from_terms(nf, list(NewOperator("Lz", nf, up, dn).terms.items()))
    Sz = dense_from_terms(nf, list(NewOperator("Sz", nf, up, dn).terms.items()))
    N = dense_from_terms(nf, list(NewOperator("Number", nf, up + dn, up + dn, [1] * 10).terms.items()))
    assert np.allclose(Lz @ Sz, Sz @ Lz)
    assert np.allclose(np.diag(N), [bin(i).count("1") for i in range(2 ** nf)])
    ls = dense_from_terms(nf, list(NewOperator("ldots", nf, up, dn).terms.items()))
    # one d electron: l.s eigenvalues l/2 = 1 (6-fold) and -(l+1)/2 = -1.5 (4-fold)
    one = [i for i in range(2 ** nf) if bin(i).count("1") == 1]
    w = np.linalg.eigvalsh(ls[np.ix_(one, one)])
    assert np.allclose(np.sort(w), [-1.5] * 4 + [1.0] * 6)


def test_wavefunction_and_options_errors():
    psi = NewWavefunction(4, 0, [["1100", 1], ["0011", 1j], ["1100", 1]])
    w, c = psi.host()
    assert len(w) == 2 and c[list(map(int, w[:, 0])).index(max(map(int, w[:, 0])))] == 2
    with pytest.raises(QSError):
        NewWavefunction(4, 1, [["1100", 1]])
    with pytest.raises(QSError):
        NewOperator(3, 0, [[3, -0.0, 1]])
    from quanty_amd.api import _options
    with pytest.raises(QSError):
        _options([["Emim", 1]], {"Emin": 0})


def test_conserved_groups():
    from quanty_amd import problems as P
    from quanty_amd.api import conserved_groups
    ni = P.ni_xas()
    gi = conserved_groups(Operator(26, ni["Hi"]))
    # Hi never touches the 2p modes; d and ligand exchange electrons
    assert all([m] in gi for m in range(6))
    assert list(range(6, 26)) in gi
    gf = conserved_groups(Operator(26, ni["Hf"]))
    # Hf mixes the 2p modes (spin-orbit) and keeps pd exchange count-preserving
    assert list(range(0, 6)) in gf and list(range(6, 26)) in gf
    ch = conserved_groups(NewOperator(4, 0, [[0, -1.0, 1], [1, -0.0, 1], [2, 3, -2.0, -3.0, 1]]))
    assert [0, 1] in ch and [2] in ch and [3] in ch


def test_spectra_width_tables_match_oracle_interpolation():
    """Spectra.Broaden's per-point widths from {{E, FWHM}, ...} tables
    (LuaInterface/Spectra.cpp:37-62), against the oracle's restatement."""
    import oracle
    from quanty_amd.api import Spectra

    S = Spectra(-13.5, 21.5, 3500, 0.1, np.zeros((1, 3501), dtype=complex))
    dE = 35.0 / 3500
    for t in ([[-13.5, 0.46], [3, 0.46], [3.01, 0.8], [21.5, 0.9]], [[0.0, 0.3]], [[-20, 1], [-15, 2]],
              [[30, 1], [40, 2]], [[-1, 0.2], [-1, 0.4], [2, 0.1]]):
        got = S._widths(t, "Gaussian")
        ref = oracle.fwhm_table(t, -13.5, dE, 3501)
        assert np.allclose(got, ref, rtol=0, atol=1e-15), t
    assert S._widths(0.5, "Gaussian") == 0.5
    with pytest.raises(QSError):
        S._widths([[1, 0.2], [0, 0.3]], "Gaussian")
    with pytest.raises(QSError):
        S._widths([[1, 0.2, 3]], "Lorentzian")


def test_named_angular_momentum_operators_match_dense():
    """Named shell operators (LuaInterface/Operator.cpp named forms) on a p shell:
    S^2, L^2, J^2 have eigenvalues x(x+1); S+ = Sx + i Sy, J = L + S, and
    [Jz, J+] = J+, checked on the Jordan-Wigner matrices of 6 modes."""
    nf = 6
    up, dn = [1, 3, 5], [0, 2, 4]
    M = {n: dense_from_terms(nf, list(NewOperator(n, nf, up, dn).terms.items()))
         for n in ("Sx", "Sy", "Sz", "Splus", "Smin", "Ssqr", "Lx", "Ly", "Lz", "Lplus", "Lmin", "Lsqr",
                   "Jx", "Jy", "Jz", "Jplus", "Jmin", "Jsqr")}
    assert np.allclose(M["Splus"], M["Sx"] + 1j * M["Sy"]) and np.allclose(M["Smin"], M["Sx"] - 1j * M["Sy"])
    assert np.allclose(M["Lplus"], M["Lx"] + 1j * M["Ly"]) and np.allclose(M["Jmin"], M["Jx"] - 1j * M["Jy"])
    for k in "xyz":
        assert np.allclose(M["J" + k], M["L" + k] + M["S" + k])
    assert np.allclose(M["Jz"] @ M["Jplus"] - M["Jplus"] @ M["Jz"], M["Jplus"])
    assert np.allclose(M["Lx"] @ M["Ly"] - M["Ly"] @ M["Lx"], 1j * M["Lz"])
    for sq, allowed in (("Ssqr", [s * (s + 1) for s in np.arange(0, 3.5, 0.5)]),
                        ("Lsqr", [l * (l + 1) for l in range(0, 5)]),
                        ("Jsqr", [j * (j + 1) for j in np.arange(0, 6.5, 0.5)])):
        ev = np.linalg.eigvalsh(M[sq])
        assert all(np.min(np.abs(np.array(allowed) - e)) < 1e-10 for e in ev)
    # two electrons in a p shell: S = 0, 1 and L = 0, 1, 2 all occur
    idx = [i for i in range(2 ** nf) if bin(i).count("1") == 2]
    ev = np.linalg.eigvalsh(M["Lsqr"][np.ix_(idx, idx)])
    assert set(np.round(ev).astype(int)) == {0, 2, 6}


@pytest.mark.parametrize("energies", [(0.6, -0.4), (1.0, 0.0), (0.0, 1.0), (-0.3, 2.2)])
def test_potential_oh_d_equals_eg_t2g_projectors(energies):
    """NewOperator("CF", ..., PotentialExpandedOnClm("Oh", 2, {e_eg, e_t2g})) is
    e_eg P_eg + e_t2g P_t2g on the d shell (the cubic harmonics eg = {z^2,
    x^2-y^2}, t2g = {xy, yz, zx}), and between two shells the same projector
    coupling (TestScripts/Eigensystem.Quanty:894-930 builds 10Dq, the eg/t2g
    counts and the hybridization this way)."""
    from quanty_amd import operators as O
    from quanty_amd.api import PotentialExpandedOnClm
    akm = PotentialExpandedOnClm("Oh", 2, list(energies))
    up, dn = [1, 3, 5, 7, 9], [0, 2, 4, 6, 8]
    upl, dnl = [11, 13, 15, 17, 19], [10, 12, 14, 16, 18]
    eg, t2g = O.oh_d_projectors()
    M = O.d_shell_matrix_to_spin(energies[0] * eg + energies[1] * t2g)
    md, ml = O.shell_modes(up, dn), O.shell_modes(upl, dnl)
    for got, want in ((NewOperator("CF", 20, up, dn, akm), O.one_body(md, md, M)),
                      (NewOperator("CF", 20, up, dn, upl, dnl, akm), O.one_body(md, ml, M))):
        diff = got - Operator(20, want)
        assert max([abs(v) for v in diff.terms.values()] or [0.0]) < 1e-14
    assert PotentialExpandedOnClm("Oh", 1, [0.5]) == [[0, 0, 0.5]]
    with pytest.raises(QSError):
        PotentialExpandedOnClm("D6h", 2, [1, 2, 3, 4])


def test_mlft_hamiltonian_as_the_script_writes_it():
    """The d5 ligand-field Hamiltonian written with the named operators the way
    TestScripts/Eigensystem.Quanty:794-990 writes it (NewOperator "U", "ldots",
    "Sz", "Lz", "CF" with PotentialExpandedOnClm, "Number") equals the model
    whose eigenstates reproduce the script's expectation-value table."""
    from models import mlft_d5
    from quanty_amd.api import PotentialExpandedOnClm
    NF = 20
    up_d, dn_d = [1, 3, 5, 7, 9], [0, 2, 4, 6, 8]
    up_l, dn_l = [11, 13, 15, 17, 19], [10, 12, 14, 16, 18]
    nd, U, Delta = 5, 5.5, 8.0
    F2, F4, tenDq, tenDqL, Veg, Vt2g, zeta, B = 9.35, 5.78, 0.67, 1.06, 1.92, 1.15, 0.04, 1e-6
    ed = (10 * Delta - nd * (19 + nd) * U / 2) / (10 + nd)
    eL = nd * ((1 + nd) * U / 2 - Delta) / (10 + nd)
    F0 = U + (F2 + F4) * 2 / 63
    akm = PotentialExpandedOnClm("Oh", 2, [0.6, -0.4])
    eg = PotentialExpandedOnClm("Oh", 2, [1, 0])
    t2g = PotentialExpandedOnClm("Oh", 2, [0, 1])
    N_3d = NewOperator("Number", NF, up_d, up_d, [1] * 5) + NewOperator("Number", NF, dn_d, dn_d, [1] * 5)
    N_L = NewOperator("Number", NF, up_l, up_l, [1] * 5) + NewOperator("Number", NF, dn_l, dn_l, [1] * 5)
    H = (F0 * NewOperator("U", NF, up_d, dn_d, [1, 0, 0]) + F2 * NewOperator("U", NF, up_d, dn_d, [0, 1, 0])
         + F4 * NewOperator("U", NF, up_d, dn_d, [0, 0, 1]) + zeta * NewOperator("ldots", NF, up_d, dn_d)
         + B * (2 * NewOperator("Sz", NF, up_d, dn_d) + NewOperator("Lz", NF, up_d, dn_d))
         + tenDq * NewOperator("CF", NF, up_d, dn_d, akm) + tenDqL * NewOperator("CF", NF, up_l, dn_l, akm)
         + Veg * (NewOperator("CF", NF, up_d, dn_d, up_l, dn_l, eg) + NewOperator("CF", NF, up_l, dn_l, up_d, dn_d, eg))
         + Vt2g * (NewOperator("CF", NF, up_d, dn_d, up_l, dn_l, t2g)
                   + NewOperator("CF", NF, up_l, dn_l, up_d, dn_d, t2g))
         + ed * N_3d + eL * N_L)
    _, terms, _, _ = mlft_d5("z")
    diff = H - Operator(NF, terms)
    assert max([abs(v) for v in diff.terms.values()] or [0.0]) < 1e-13


def test_determinant_string():
    """TestScripts/DeterminantString.Quanty with the script's index lists spelled
    out (22 modes: H 1s = 0-1, Fe 2s = 2-3, Fe 2p = 4-9, Fe 3s = 10-11,
    Fe 3d = 12-21)."""
    from quanty_amd.api import DeterminantString
    NF = 22
    H_1s, Fe_2p, Fe_3d = [0, 1], list(range(4, 10)), list(range(12, 22))
    assert DeterminantString(NF, [0, 1, 2, 3]) == "1111000000000000000000"
    assert DeterminantString(NF, Fe_2p) == "0000111111000000000000"
    assert DeterminantString(NF, H_1s) == "1100000000000000000000"
    assert DeterminantString(NF, list(range(2, 22))) == "0011111111111111111111"
    assert DeterminantString(NF, H_1s, Fe_2p, Fe_3d) == "1100111111001111111111"
    psi = NewWavefunction(NF, 0, [[DeterminantString(NF, Fe_3d), 1]])
    assert psi == NewWavefunction(22, 0, [["0000000000001111111111", 1]])
    with pytest.warns(UserWarning):
        assert DeterminantString(4, [1, 1, 7]) == "0100"


def test_bitmask_index_codec():
    """TestScripts/BitMaskToIndex.Quanty and IndexToBitMask.Quanty."""
    from quanty_amd.api import BitMaskToIndex, IndexToBitMask
    for m in ("00000 10101 000", "0000010101000", "  0000010101000", "0000010101000   "):
        assert BitMaskToIndex(m) == (13, [5, 7, 9])
    assert IndexToBitMask(15, [0, 1, 2]) == "111000000000000"
    assert IndexToBitMask(15, [0, 1, 2, 14]) == "111000000000001"
    # the reference's error behaviour (LuaIndexToBitMask, LuaInterface/ssp.cpp:6050-6053)
    # and BitMaskToIndex skipping every character but 0 and 1 (Core/Utilities.cpp:5-31)
    with pytest.raises(QSError, match="smaller than zero"):
        IndexToBitMask(15, [0, -1])
    with pytest.raises(QSError, match="larger or equal than NF"):
        IndexToBitMask(15, [3, 15])
    with pytest.raises(QSError, match="negative value for NF"):
        IndexToBitMask(-1, [])
    assert IndexToBitMask(4, []) == "0000"
    assert BitMaskToIndex("00a1-1|0x") == (5, [2, 3])
    assert BitMaskToIndex("") == (0, [])


def test_count_saturates_instead_of_wrapping():
    """Sector counts past 2^63 (68 modes holding 34, 76 modes holding 38) saturate
    at INT64_MAX instead of wrapping to negative numbers; enumeration refuses
    sectors of 2^32 - 1 or more determinants with a clear error."""
    from quanty_amd.engine import Basis
    for nf in (68, 76):
        n = Basis.count(nf, [("1" * nf, nf // 2, nf // 2)])
        assert n == 2 ** 63 - 1
    # exact below the limit: C(64, 32) fits in int64
    from math import comb
    assert Basis.count(64, [("1" * 64, 32, 32)]) == comb(64, 32)
    assert Basis.count(40, [("1" * 40, 20, 20)]) == comb(40, 20)


def test_normal_order_against_jordan_wigner():
    """The canonical normal order of random ladder strings (lengths 1-4, 5
    modes) equals the string as a Jordan-Wigner matrix; strings that only
    reorder one product give identical canonical terms."""
    from models import dense_from_terms, jw_matrices
    from quanty_amd import operators as Ops
    rng = np.random.default_rng(3)
    nf = 5
    cs = jw_matrices(nf)
    for _ in range(40):
        terms = []
        for _ in range(4):
            ln = int(rng.integers(1, 5))
            codes = tuple(int(rng.integers(0, nf)) | (Ops.CR if rng.random() < 0.5 else 0) for _ in range(ln))
            terms.append((codes, complex(rng.normal(), rng.normal())))
        no = Ops.normal_order(terms)
        assert np.allclose(dense_from_terms(nf, list(no.items()), cs), dense_from_terms(nf, terms, cs), atol=1e-12)
        for c in no:  # canonical: strictly ascending codes
            assert all(a < b for a, b in zip(c, c[1:]))
    # c+_0 c_1 = - c_1 c+_0
    a = Ops.normal_order([((Ops.an(1), Ops.cr(0)), 1.0)])
    b = Ops.normal_order([((Ops.cr(0), Ops.an(1)), -1.0)])
    assert a == b


def test_hermiticity_classes_and_repair():
    """QOperatorHermitian's three answers (Core/OperatorArithmetic.cpp:6517-6568)
    and the almost-Hermitian repair of LuaEigensystem (ssp.cpp:4164-4205);
    CreateSpectra and Eigensystem refuse a non-Hermitian operator before any
    device work."""
    from quanty_amd.api import CreateSpectra, Eigensystem, hermitian_class, hermitian_part
    nf = 4
    H = NewOperator(nf, 0, [[0, -1, 0.5 + 0.2j], [1, -0.0, 0.5 - 0.2j], [0, -0.0, 1.0], [2, 3, -3, -2, 2.0]])
    assert hermitian_class(H, 1e-6) == 1
    # the same operator written in another order is still Hermitian
    H2 = NewOperator(nf, 0, [[0, -1, 0.5 + 0.2j], [-0.0, 1, -0.5 + 0.2j], [0, -0.0, 1.0], [3, 2, -2, -3, 2.0]])
    assert hermitian_class(H2, 1e-6) == 1
    almost = H + NewOperator(nf, 0, [[1, -2, 1e-9]])
    assert hermitian_class(almost, 1e-6) == 2
    fixed = hermitian_part(almost, 1e-6)
    assert hermitian_class(fixed, 1e-6) == 1
    assert abs(fixed.terms[(2, 0x8001)] - 0.5e-9) < 1e-20
    bad = H + NewOperator(nf, 0, [[1, -2, 1e-3]])
    assert hermitian_class(bad, 1e-6) == 0
    psi = NewWavefunction(nf, 0, [["1100", 1.0]])
    with pytest.raises(QSError, match="not Hermitian"):
        Eigensystem(bad, [nf, 0, ["1111", 2, 2]], 1)
    with pytest.raises(QSError, match="not Hermitian"):
        CreateSpectra(bad, NewOperator("An", nf, 0), psi)


def test_option_names_of_the_reference():
    """Eigensystem and CreateSpectra take every option name of LuaEigensystem /
    LuaCreateSpectra (ssp.cpp:4062-4078, 1958-1981) that this engine covers,
    reject unknown ones and name the unsupported ones, before any device work."""
    from quanty_amd.api import (_CS_OPTIONS, _CS_UNSUPPORTED, _EIG_ALIASES, _EIG_OPTIONS, CreateSpectra,
                                Eigensystem, _options)
    for name in ["DenseBorder", "NKrylovStart", "NKrylovMax", "NKrylovStep", "NKrylovMin", "NBitsKey", "epsilon",
                 "Epsilon", "SingularValue", "zero", "Zero", "restrictions", "Restrictions", "ExpandBasis",
                 "ShiftScheme"]:
        assert name in _EIG_OPTIONS or name in _EIG_ALIASES
    o = _options([["zero", 1e-9], ["Zero", 1e-10]], _EIG_OPTIONS, _EIG_ALIASES)
    assert o["Zero"] == 1e-10  # both spellings write one setting; the later pair wins
    o = _options([["restrictions", [4, 0, ["1100", 1, 2]]]], _EIG_OPTIONS, _EIG_ALIASES)
    assert o["Restrictions"] == [4, 0, ["1100", 1, 2]]
    nf = 4
    H = NewOperator(nf, 0, [[0, -0.0, 1.0]])
    with pytest.raises(QSError, match="unknown option"):
        Eigensystem(H, [nf, 0, ["1111", 2, 2]], 1, [["NoSuchOption", 1]])
    with pytest.raises(QSError, match="ShiftScheme"):
        Eigensystem(H, [nf, 0, ["1111", 2, 2]], 1, [["ShiftScheme", "Sideways"]])
    psi = NewWavefunction(nf, 0, [["1100", 1.0]])
    for name in _CS_UNSUPPORTED:
        with pytest.raises(QSError, match="not supported"):
            CreateSpectra(H, NewOperator("An", nf, 0), psi, [[name, 1]])
    for name in ["Tensor", "KrylovBasis", "NTri", "epsilon", "Epsilon", "SingularValue", "restrictions",
                 "Restrictions", "Emin", "Emax", "NE", "Gamma", "E0", "DenseBorder", "ForceDenseMethod", "Method"]:
        assert name in _CS_OPTIONS or name in ("epsilon", "restrictions")
    many = [NewOperator("An", nf, i % nf) for i in range(17)]
    with pytest.raises(QSError, match="at most 16"):
        CreateSpectra(H, many, psi, [["Tensor", True]])


def _akm(lst):
    return {(k, m): complex(a) for k, m, a in lst}


def test_crystal_field_tables_against_the_reference():
    """PotentialExpandedOnClm from projectors on symmetry-adapted functions
    equals the reference's closed forms (Core/BasicMath/BasicMath_StandardFunctions.cpp:
    Oh :4648, Td :4322, D4h :3213 (l = 1 and 2), C4v :1832, D2h :2899, D3d :3563)
    for random energies."""
    import math
    from quanty_amd.api import PotentialExpandedOnClm
    rng = np.random.default_rng(5)
    s = math.sqrt
    for _ in range(5):
        E = rng.normal(size=6)
        want = {
            ("Oh", 2): {(0, 0): 0.4 * E[0] + 0.6 * E[1], (4, 0): 2.1 * (E[0] - E[1]),
                        (4, 4): 1.5 * s(0.7) * (E[0] - E[1]), (4, -4): 1.5 * s(0.7) * (E[0] - E[1])},
            ("D4h", 1): {(0, 0): (E[0] + 2 * E[1]) / 3, (2, 0): 5 * (E[0] - E[1]) / 3},
            ("D4h", 2): {(0, 0): (E[0] + E[1] + E[2] + 2 * E[3]) / 5, (2, 0): E[0] - E[1] - E[2] + E[3],
                         (4, 0): 0.3 * (6 * E[0] + E[1] + E[2] - 8 * E[3]),
                         (4, 4): 0.15 * s(70) * (E[1] - E[2]), (4, -4): 0.15 * s(70) * (E[1] - E[2])},
            ("D2h", 2): {(0, 0): 0.2 * (E[0] + E[1] + E[3] + E[4] + E[5]),
                         (2, 0): 0.5 * (-2 * E[0] + 2 * E[1] - 2 * E[3] + E[4] + E[5]),
                         (4, 0): 0.3 * (E[0] + 6 * E[1] + E[3] - 4 * E[4] - 4 * E[5]),
                         (2, 2): 0.25 * (s(6) * (E[4] - E[5]) - 4 * s(2) * E[2]),
                         (2, -2): 0.25 * (s(6) * (E[4] - E[5]) - 4 * s(2) * E[2]),
                         (4, 2): s(0.9) * (E[4] - E[5] + s(3) * E[2]), (4, -2): s(0.9) * (E[4] - E[5] + s(3) * E[2]),
                         (4, 4): 1.5 * s(0.7) * (E[0] - E[3]), (4, -4): 1.5 * s(0.7) * (E[0] - E[3])},
        }
        want[("Td", 2)] = want[("O", 2)] = want[("Oh", 2)]
        want[("C4v", 2)] = want[("D4h", 2)]
        r3 = (2 * s(35) * E[1] - 2 * s(35) * E[2] - s(70) * E[3]) / 10
        want[("D3d", 2)] = {(0, 0): (E[0] + 2 * E[1] + 2 * E[2]) / 5, (2, 0): E[0] - E[1] - 2 * s(2) * E[3],
                            (4, 0): (9 * E[0] - 2 * E[1] - 7 * E[2] + 10 * s(2) * E[3]) / 5,
                            (4, -3): complex(r3, -r3), (4, 3): complex(-r3, -r3)}
        for (g, l), ref in want.items():
            got = _akm(PotentialExpandedOnClm(g, l, list(E)))
            for key in set(got) | set(ref):
                assert abs(got.get(key, 0) - ref.get(key, 0)) < 1e-12, (g, l, key)
    with pytest.raises(QSError, match="not found"):
        PotentialExpandedOnClm("D6h", 2, [1, 2, 3])


def test_crystal_field_levels_on_one_electron():
    """The CF operator built from PotentialExpandedOnClm gives a single d
    electron exactly the requested levels with the representation
    degeneracies (times 2 for spin), for tetragonal, orthorhombic (no mixing)
    and trigonal symmetry."""
    from models import dense_from_terms
    from quanty_amd.api import NewOperator, PotentialExpandedOnClm
    up, dn = [1, 3, 5, 7, 9], [0, 2, 4, 6, 8]
    for g, E, levels in [("D4h", [0.3, 1.1, -0.7, 0.2], [0.3, 1.1, -0.7, 0.2, 0.2]),
                         ("D2h", [0.3, 1.1, 0.0, -0.7, 0.2, 0.5], [0.3, 1.1, -0.7, 0.2, 0.5]),
                         ("D3d", [0.3, 1.1, -0.7, 0.0], [0.3, 1.1, 1.1, -0.7, -0.7]),
                         ("Oh", [0.6, -0.4], [0.6, 0.6, -0.4, -0.4, -0.4])]:
        op = NewOperator("CF", 10, up, dn, PotentialExpandedOnClm(g, 2, E))
        Hd = dense_from_terms(10, op.term_list())
        one = [1 << (9 - i) for i in range(10)]  # the ten one-electron determinants
        w = np.linalg.eigvalsh(Hd[np.ix_(one, one)])
        assert np.allclose(w, np.sort(np.repeat(levels, 2)), atol=1e-12), g
