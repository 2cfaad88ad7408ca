"""`.Quanty` scripts through the Lua front end on the GPU, checked against the
oracle and the reference's golden vectors:

* tests/lua/ni2p3d_xas.Quanty, a Ni2+ L2,3 XAS script in the style of the
  reference's tutorials: its three lowest energies equal the lowest
  eigenvalues of the oracle's dense H on the 2p6 3d8 sector, the printed
  expectation values equal the oracle's <psi|O|psi>, the spectra it writes
  with Print equal the oracle's Lanczos + continued fraction, broadened by
  the oracle's Gaussian and energy-dependent Lorentzian, and its Lanczos
  branch equals the same;
* the impurity + chain model of the reference's CreateSpectra test written in
  Lua (raw operator tables with -0, loops, I): the Tri tables equal the golden
  ones (TestScripts/CreateSpectra.Quanty:188-260);
* `O*psi` from Lua gives the reference's OpTimesPsi answer, and the two
  Broaden forms behave as the reference's (in place / copy)."""
import io
import os

import numpy as np
import pytest

import oracle
from models import golden
from quanty_amd import api
from quanty_amd import det as D
from quanty_amd import lua

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def dense(nf, terms, dets):
    """The matrix of an operator on a list of determinants, column by column
    from the oracle's H*psi."""
    op = oracle.Op(nf, terms)
    index = {bytes(r): i for i, r in enumerate(dets)}
    M = np.zeros((len(dets), len(dets)), dtype=complex)
    for j in range(len(dets)):
        d, v = op.apply(oracle.Vec.from_arrays(nf, dets[j:j + 1], np.ones(1, dtype=complex))).arrays()
        for row, val in zip(d, v):
            M[index[bytes(row)], j] += val
    return M


def read_spectra(path):
    """(energies, complex spectra (nspec, npts)) of a file written by Print."""
    rows = []
    with open(path) as f:
        lines = f.read().splitlines()
    assert lines[0].startswith("#Spectra:")
    nspec = int(lines[0].split()[1])
    for ln in lines[5:]:
        rows.append([float(x) for x in ln.split()])
    A = np.array(rows)
    assert A.shape[1] == 1 + 2 * nspec
    return A[:, 0], (A[:, 1::2] + 1j * A[:, 2::2]).T


def oracle_vec(psi):
    w, c = psi.host()
    return oracle.Vec.from_arrays(psi.NF, D.words_to_bytes(w, psi.NF), c)


def test_ni_xas_script_against_the_oracle(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    out = io.StringIO()
    it = lua.run_file(os.path.join(HERE, "lua", "ni2p3d_xas.Quanty"), out=out)
    text = out.getvalue().splitlines()
    G = it.G.get
    nf = 16
    Hi, Hf = G("H_i"), G("H_f")
    psis = G("psiList").array()
    assert len(psis) == 3
    # energies: the three lowest eigenvalues of the oracle's dense H_i on 2p6 3d8
    sector = oracle.enumerate_sector(nf, [("111111" + "0" * 10, 6, 6), ("0" * 6 + "1" * 10, 8, 8)])
    assert len(sector) == 45
    Mi = dense(nf, Hi.term_list(), sector)
    assert np.max(np.abs(Mi - Mi.conj().T)) < 1e-12
    ev = np.linalg.eigvalsh(Mi)[:3]
    # the printed table: <E>, <S^2>, ... of each state, equal to the oracle's expectation values
    ops = G("ops").array()
    table = [ln.split() for ln in text if ln[:3].strip().isdigit()]
    assert len(table) == 3
    hi_op = oracle.Op(nf, Hi.term_list())
    for i, psi in enumerate(psis):
        v = oracle_vec(psi)
        e = v.dot(hi_op.apply(v)).real / v.dot(v).real
        assert abs(e - ev[i]) < 1e-9 * max(1.0, abs(ev[i]))
        for j, O in enumerate(ops):
            want = v.dot(oracle.Op(nf, O.term_list()).apply(v)).real
            assert abs(float(table[i][j + 1]) - want) <= 6e-6, (i, j)
        assert abs(float(table[i][9]) - 8.0) < 1e-5  # <Nd>
    assert "spectra:\t9\tpoints:\t1501" in text
    # the spectra written by XAS.Print after XAS.Broaden(0.3, {{-5, 0.4}, {5, 0.8}})
    E, got = read_spectra(tmp_path / "ni_xas.dat")
    assert got.shape == (9, 1501) and abs(E[0] + 10) < 1e-12 and abs(E[-1] - 20) < 1e-9
    dE = 30.0 / 1500
    hf_op = oracle.Op(nf, Hf.term_list())
    raw = []
    for T in (G("Tz"), G("Tr"), G("Tl")):
        for psi in psis:
            v = oracle_vec(psi)
            E0 = v.dot(hf_op.apply(v)).real
            tv = oracle.Op(nf, T.term_list()).apply(v)
            a, b = hf_op.lanczos(tv, 60)
            raw.append(oracle.cf(a, b, E0, -10.0, 20.0, 1500, 0.2))
    raw = np.array(raw)
    want = oracle.broaden(raw, dE, "G", 0.3)
    want = oracle.broaden(want, dE, "L", oracle.fwhm_table([[-5, 0.4], [5, 0.8]], -10.0, dE, 1501), per_point=True)
    scale = np.max(np.abs(want))
    assert np.max(np.abs(got - want)) <= 1e-8 * scale
    # the isotropic sum over the ground state's three polarizations
    _, iso = read_spectra(tmp_path / "ni_xas_iso.dat")
    assert np.max(np.abs(iso[0] - (want[0] + want[3] + want[6]))) <= 1e-8 * scale
    assert np.all(iso[0].imag <= 1e-12 * scale)  # absorption: -Im G >= 0
    # the Lanczos branch (60 steps) on the ground state, unbroadened
    assert "tridiagonal:\t1\tTri\t61\t60" in text
    _, lz = read_spectra(tmp_path / "ni_xas_lanczos.dat")
    assert np.max(np.abs(lz[0] - raw[0])) <= 1e-8 * np.max(np.abs(raw[0]))


CHAIN = """
NF = 2 + 2 * NBath
U, t = 3.0, 1.0
h = (kind == 0) and t or I * t
if kind == 2 then H = NewOperator(NF, 0, {}) else H = NewOperator(NF, 0, {{0, -0, 1, -1, U}}) end
for i = 1, NBath do
  H = H + NewOperator(NF, 0, {{2 * i, -2 * (i - 1), h}, {2 * i + 1, -(2 * (i - 1) + 1), h},
                              {2 * (i - 1), -2 * i, Conjugate(h)}, {2 * (i - 1) + 1, -(2 * i + 1), Conjugate(h)}})
end
T = NewOperator("Cr", NF, spin)
zeros = string.rep("0", NF - 2)
if which == 1 then
  psi = NewWavefunction(NF, 0, {{"10" .. zeros, 1 / math.sqrt(2)}, {"01" .. zeros, 1 / math.sqrt(2)}})
else
  psi = NewWavefunction(NF, 0, {{"10" .. zeros, 1 / math.sqrt(5)}, {"01" .. zeros, -2 * I / math.sqrt(5)}})
end
opts = {{"DenseBorder", 1}, {"Tensor", false}}
if restricted then
  opts[3] = {"Restrictions", {NF, 0, {"00" .. string.rep("0", NF - 4) .. "11", 0, 0}}}
end
spec, G = CreateSpectra(H, T, psi, opts)
"""


@pytest.mark.parametrize("name", ["G0Dn1", "G1Up2", "G2Dn1", "G2Up2"])
@pytest.mark.parametrize("block", [1, 2])
def test_chain_spectra_from_lua_equal_the_golden_tables(name, block):
    g = golden("createspectra_tri.json")
    ref = g["tables"][name]
    it = lua.Interpreter(io.StringIO())
    for k, v in {"NBath": 5.0 if block == 1 else 6.0, "kind": float(name[1]),
                 "spin": 0.0 if name[2:4] == "Dn" else 1.0, "which": float(name[-1]),
                 "restricted": block == 2}.items():
        it.G.set(k, v)
    it.run(CHAIN, "chain")
    G = it.G.get("G")
    tri = G.get(1)
    assert tri.get("type") == "Tri" and tri.get("mu") == 0
    a = np.array(lua.to_py(tri.get(1)), dtype=float)
    b = np.array(lua.to_py(tri.get(2)), dtype=float)
    assert np.allclose(a[:len(ref["a"])], ref["a"], atol=5e-13, rtol=1e-12)
    assert np.allclose(b[:len(ref["b"])], ref["b"], atol=5e-13, rtol=1e-12)
    thr = g["threshold_block1"] if block == 1 else g["threshold_block2"]
    for w in g["omega"]:
        want = oracle.cf(np.array(ref["a"]), np.array(ref["b"]), 0.0, w, w + 1.0, 1, g["gamma"])[0]
        got = it.call(tri, [w, g["gamma"]])[0]  # the table is callable as G(omega, Gamma)
        assert abs(got - want) < thr


def test_operator_times_psi_from_lua():
    """O = i c+_0 c_0 + c+_0 c+_1 c_0 c_1 on |11>: O psi = (-1 + i)|11> (the
    sign known-answer of the reference's OpTimesPsi example)."""
    out = io.StringIO()
    it = lua.Interpreter(out)
    it.run("""
O = NewOperator(2, 0, {{0, -0, I}, {0, 1, -0, -1, 1}})
psi = NewWavefunction(2, 0, {{"11", 1}})
Opsi = O * psi
print(Braket(psi, Opsi), psi * O * psi, Opsi.N)
""")
    assert out.getvalue() == "(-1 + 1 I)\t(-1 + 1 I)\t1\n"


def test_broaden_forms_from_lua():
    G = np.zeros((2, 101), dtype=complex)
    G[0, 50] = -1j
    G[1, 30] = -2j
    it = lua.Interpreter(io.StringIO())
    it.G.set("S", api.Spectra(-5.0, 5.0, 100, 0.1, G.copy()))
    it.run("""
B = Spectra.Broaden(S, 0.5, {{-1, 0.2}, {1, 0.6}})
untouched = S.Copy()
S.Broaden(0.5, {{-1, 0.2}, {1, 0.6}})
""")
    want = api.Spectra(-5.0, 5.0, 100, 0.1, G.copy()).Broaden(0.5, [[-1, 0.2], [1, 0.6]]).G
    assert np.allclose(it.G.get("B").G, want, rtol=0, atol=1e-14)
    assert np.allclose(it.G.get("S").G, want, rtol=0, atol=1e-14)
    assert np.array_equal(it.G.get("untouched").G, G)
    with pytest.raises(lua.LuaError, match="2 arguments expected"):
        it.run("S.Broaden(0.5)")
