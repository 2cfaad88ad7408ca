"""Object text formats (io.put / io.get, Spectra.Print) without a GPU: number
formatting against the C library's printf (the reference writes with fprintf),
the layout of the reference's writers (LuaInterface/lua_extensions/get_put.cpp:503-632,
IO/StreamPrint.cpp:4-60), and exact put -> get round trips."""
import ctypes
import ctypes.util

import numpy as np
import pytest

from quanty_amd import QSError, textio
from quanty_amd.api import NewOperator, NewWavefunction, Spectra

_libc = ctypes.CDLL(ctypes.util.find_library("c"))


def cfmt(fmt: str, *args) -> str:
    buf = ctypes.create_string_buffer(256)
    cargs = [ctypes.c_double(a) if isinstance(a, float) else ctypes.c_int(a) for a in args]
    _libc.snprintf(buf, 256, fmt.encode(), *cargs)
    return buf.value.decode()


@pytest.mark.parametrize("x", [0.0, -0.0, 1.0, -1.0 / 3, 2.5e-300, 6.02214076e23, 1e-5, 123456.789, -9.999999999e99])
def test_number_formats_match_c_printf(x):
    assert "%23.16E" % x == cfmt("%23.16E", x)
    assert " %22.15E" % x == cfmt(" %22.15E", x)
    assert "%19.12E" % x == cfmt("%19.12E", x)
    assert "%c%4u," % ("-", 7) == cfmt("%c%4u,", ord("-"), 7)


def test_operator_text_layout_and_round_trip(tmp_path):
    O = NewOperator(10, 0, [[2, -0.0, 0.5], [0, -0.0, 1], [3, 1, -2, -5, 0.25 - 1e-3j], [7, 1 / 3]],
                    [["Name", "H"], ["Restrictions", [10, 0, ["11 0000 0000", 1, 2]]]])
    t = textio.format_operator(O)
    lines = t.split("\n")
    assert lines[0] == "NewOperator( 10, 0,{"
    # real terms: "%c%4u," per ladder operator, leftmost written first
    assert lines[2] == " {" + cfmt("%c%4u,", ord(" "), 2) + cfmt("%c%4u,", ord("-"), 0) + cfmt("%23.16E", 0.5) + " },"
    # complex length-4 term: " %c%4u," then "re - I *|im| }"
    assert lines[4].startswith(" {     3,     1, -   2, -   5, 2.5000000000000000E-01 - I * 1.0000000000000000E-03 }")
    assert '{"Restrictions",{ 10, 0, {"1100000000",1,2}}}' in t
    path = textio.put(str(tmp_path), "H", O)
    assert open(path).read().startswith("H = NewOperator( 10, 0,{")
    back = textio.get(str(tmp_path), "H")
    assert back.terms == O.terms and back.Name == "H"
    with pytest.raises(QSError):
        textio.put(str(tmp_path), "not valid", O)


def test_wavefunction_text_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    strings = sorted({"".join(rng.choice(["0", "1"], 70)) for _ in range(20)})
    psi = NewWavefunction(70, 0, [(s, complex(*rng.normal(size=2))) for s in strings])
    t = textio.format_wavefunction(psi)
    assert t.split("\n")[0] == "NewWavefunction( 70, 0,{"
    assert t.split("\n")[1].startswith('{"' + strings[0] + '",')
    textio.put(str(tmp_path), "psi", psi)
    back = textio.get(str(tmp_path), "psi")
    w1, c1 = psi.host()
    w2, c2 = back.host()
    assert np.array_equal(w1, w2) and np.array_equal(c1, c2)


def test_spectra_print_layout(tmp_path):
    G = np.arange(12).reshape(2, 6) * (1 - 0.5j)
    S = Spectra(-1.0, 1.5, 5, 0.1, G, -0.5, 0.75)
    text = S.Print(str(tmp_path / "s.dat"))
    assert open(tmp_path / "s.dat").read() == text
    lines = text.split("\n")
    assert lines[0] == "#Spectra: 2"
    assert lines[1] == "Emin______Emax     " + 2 * (cfmt(" %22.15E", -1.0) + cfmt(" %22.15E", 1.5))
    assert lines[2] == "EminPole__EmaxPole " + 2 * (cfmt(" %22.15E", -0.5) + cfmt(" %22.15E", 0.75))
    assert lines[3] == "dE________Gamma    " + 2 * (cfmt(" %22.15E", 0.5) + cfmt(" %22.15E", 0.1))
    assert lines[4] == "Energy             " + "  Re[0]" + " " * 18 + "Im[0]" + " " * 16 + "  Re[1]" + " " * 18 + "Im[1]" + " " * 16
    assert lines[5 + 3] == cfmt("%19.12E", 0.5) + "".join(cfmt(" %22.15E", g[3].real) + cfmt(" %22.15E", g[3].imag) for g in G)
    assert len(lines) == 5 + 6 + 1


def test_spectra_sum_and_arithmetic():
    rng = np.random.default_rng(1)
    G = rng.normal(size=(3, 11)) + 1j * rng.normal(size=(3, 11))
    S = Spectra(0.0, 1.0, 10, 0.1, G)
    T = S.Sum([1, 0, 2j], [0.5, 0.5, 0.5])
    assert T.G.shape == (2, 11)
    assert np.allclose(T.G[0], G[0] + 2j * G[2]) and np.allclose(T.G[1], 0.5 * G.sum(0))
    assert np.allclose((S + 2).G, G + 2) and np.allclose((2 * S).G, 2 * G) and np.allclose((S - S).G, 0)
    with pytest.raises(QSError):
        S.Sum([1, 2])


def test_spectra_shift_multiply_element_tables():
    rng = np.random.default_rng(4)
    G = rng.normal(size=(3, 21)) + 1j * rng.normal(size=(3, 21))
    S = Spectra(-1.0, 1.0, 20, 0.1, G, -0.8, 0.9)
    T = S.Shift(2.5)
    assert (T.Emin, T.Emax, T.EminPole, T.EmaxPole) == (1.5, 3.5, 1.7, 3.4) and np.array_equal(T.G, G)
    assert np.allclose(S.MultiplyByEnergy().G, G * np.linspace(-1, 1, 21)[None, :])
    assert np.array_equal(S.Element(2).G, G[1:2]) and np.array_equal(S.Element([3, 1]).G, G[[2, 0]])
    with pytest.raises(QSError):
        S.Element(4)
    with pytest.raises(QSError):
        S.Shift([1.0, 2.0, 3.0])
    back = Spectra.FromTable(S.ToTable())
    assert np.allclose(back.G, G, rtol=0, atol=0) and (back.Emin, back.Emax, back.NE) == (-1.0, 1.0, 20)
