"""Object text formats of the reference (io.put / io.get,
LuaInterface/lua_extensions/get_put.cpp:150-275, 503-640, 792-818).

put(dirname, name, obj) writes dirname/name.Quanty holding `name = <object>`
in the reference's text form, which is Lua source:
    NewOperator( NF, NB,{
     {   3,-   2, 1.0000000000000000E+00 },
     ...},
    {{"Name","..."},{"NBitsKey",{16,...}}[, {"Restrictions",{NF, NB, {"mask",min,max}, ...}}]})
    NewWavefunction( NF, NB,{
    {"0101...", 1.0000000000000000E+00},
    ...},
    {{"Name","..."},{"NBitsKey",16}})
Ladder operators are written left to right as they appear in the product
(the rightmost acts first), " i" for a creator and "-i" for an annihilator;
complex numbers as "re + I * |im|" or "re - I * |im|"; doubles with
"%23.16E", which gives every double back.

get(dirname, name) reads such a file back. It reads these two object forms
only; it does not run Lua.
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

from ._lib import QSError

_NBITSKEY = 16  # the reference's default key width (Core/StructDef.cpp:306)


def _e(x: float) -> str:
    return "%23.16E" % x


def _complex_tail(v: complex) -> str:
    return ("- I *" if v.imag < 0 else "+ I *") + _e(abs(v.imag))


def format_operator(op) -> str:
    """The io.put text of an api.Operator (get_put.cpp:503-577)."""
    nf = op.NF
    by_len: dict = {}
    for codes, v in op.terms.items():
        by_len.setdefault(len(codes), []).append((codes, complex(v)))
    maxlen = max(by_len) if by_len else 0
    parts = []
    for l in sorted(by_len):
        terms = by_len[l]
        cplx = any(v.imag != 0 for _, v in terms)
        for codes, v in terms:
            ops = ""
            for c in reversed(codes):  # the leftmost written factor acts last
                ops += (" " if cplx else "") + ("%c%4u," % (" " if c & 0x8000 else "-", c & 0x7FFF))
            if cplx:
                parts.append(ops + _e(v.real) + " " + _complex_tail(v) + " }")
            else:
                parts.append(ops + _e(v.real) + " }")
    body = ""
    for i, p in enumerate(parts):
        body += (",\n {" if i else " {") + p
    nbk = ",".join(str(_NBITSKEY) for _ in range(1, maxlen + 1))
    text = "NewOperator( %d, %d,{\n%s},\n{{\"Name\",%s},{\"NBitsKey\",{%s}}" % (
        nf, 0, body, _lua_string(op.Name or ""), nbk)
    R = op.Restrictions
    if R:
        text += ", {\"Restrictions\",{ %d, %d" % (nf, 0)
        for m, lo, hi in R[2:]:
            text += ", {\"%s\",%u,%u}" % (str(m).replace(" ", ""), int(lo), int(hi))
        text += "}}"
    return text + "})"


def format_wavefunction(psi) -> str:
    """The io.put text of an api.Wavefunction (get_put.cpp:578-632): one line
    per determinant in the stored order (ascending here)."""
    nf = psi.NF
    words, coef = psi.host()
    cplx = bool(np.any(np.asarray(coef).imag != 0))
    lines = []
    for w, c in zip(words, coef):
        s = _det_string(w, nf)
        c = complex(c)
        if cplx:
            lines.append("{\"%s\",%s %s %s}" % (s, _e(c.real), "- I *" if c.imag < 0 else "+ I *",
                                                 _e(abs(c.imag))))
        else:
            lines.append("{\"%s\",%s}" % (s, _e(c.real)))
    return "NewWavefunction( %d, %d,{\n%s},\n{{\"Name\",%s},{\"NBitsKey\",%d}})" % (
        nf, 0, ",\n".join(lines), _lua_string(""), _NBITSKEY)


def _det_string(w, nf: int) -> str:
    x = 0
    for word in np.atleast_1d(w):
        x = (x << 64) | int(word)
    total = 64 * len(np.atleast_1d(w))
    return "".join("1" if (x >> (total - 1 - i)) & 1 else "0" for i in range(nf))


def _lua_string(s: str) -> str:
    return "\"" + s.replace("\\", "\\\\").replace("\"", "\\\"").replace("\n", "\\n") + "\""


def _valid_name(name: str) -> bool:
    return re.fullmatch(r"[A-Za-z_][A-Za-z0-9_]*", name) is not None


def put(dirname: str, name: str, obj) -> str:
    """io.put: writes dirname/name.Quanty as `name = <text>`; returns the path."""
    from .api import Operator, Wavefunction

    if not _valid_name(name):
        raise QSError(f"{name} is not a valid token for a variable name (only alpha numerical characters "
                      "and '_' are allowed)")
    if isinstance(obj, Operator):
        text = format_operator(obj)
    elif isinstance(obj, Wavefunction):
        text = format_wavefunction(obj)
    else:
        raise QSError(f"io.put of a {type(obj).__name__} is not supported (Operator and Wavefunction are)")
    os.makedirs(dirname, mode=0o700, exist_ok=True)
    path = os.path.join(dirname, name + ".Quanty")
    with open(path, "w") as f:
        f.write("%s = %s" % (name, text))
    return path


# ---------------------------------------------------------------------------
# reading
# ---------------------------------------------------------------------------
_TOKEN = re.compile(r"""\s*(?:(?P<num>\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?)
                           |(?P<str>"(?:[^"\\]|\\.)*")
                           |(?P<name>[A-Za-z_][A-Za-z0-9_]*)
                           |(?P<sym>[{}(),=+\-*]))""", re.X)


def _tokens(text: str):
    pos, out = 0, []
    text = text.rstrip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise QSError(f"io.get: cannot read the text at {text[pos:pos + 30]!r}")
        pos = m.end()
        for kind in ("num", "str", "name", "sym"):
            if m.group(kind) is not None:
                out.append((kind, m.group(kind)))
                break
    return out


class _Parser:
    def __init__(self, toks):
        self.t = toks
        self.i = 0

    def peek(self, k=0):
        return self.t[self.i + k] if self.i + k < len(self.t) else (None, None)

    def take(self, want=None):
        tok = self.peek()
        if tok[0] is None or (want is not None and tok[1] != want):
            raise QSError(f"io.get: expected {want!r}, found {tok[1]!r}")
        self.i += 1
        return tok

    def value(self):
        kind, val = self.peek()
        if val == "{":
            return self.table()
        if kind == "str":
            self.i += 1
            return bytes(val[1:-1], "utf-8").decode("unicode_escape")
        return self.number()

    def number(self):
        sign = 1.0
        if self.peek()[1] == "-":
            self.take("-")
            sign = -1.0
        elif self.peek()[1] == "+":
            self.take("+")
        kind, val = self.take()
        if kind == "name" and val == "I":
            re_, im = 0.0, sign
            if self.peek()[1] == "*":
                self.take("*")
                im = sign * self._unsigned()
            return complex(re_, im)
        if kind != "num":
            raise QSError(f"io.get: expected a number, found {val!r}")
        x = math.copysign(float(val), sign)  # keeps -0 (an annihilator of mode 0)
        if self.peek()[1] in ("+", "-") and self.peek(1)[1] == "I":
            s = 1.0 if self.take()[1] == "+" else -1.0
            self.take("I")
            self.take("*")
            return complex(x, s * self._unsigned())
        return x

    def _unsigned(self):
        kind, val = self.take()
        if kind != "num":
            raise QSError(f"io.get: expected a number, found {val!r}")
        return float(val)

    def table(self):
        self.take("{")
        out = []
        while self.peek()[1] != "}":
            out.append(self.value())
            if self.peek()[1] == ",":
                self.take(",")
        self.take("}")
        return out


def parse(text: str):
    """An Operator or Wavefunction from its io.put text (with or without the
    leading `name =`)."""
    from .api import NewOperator, NewWavefunction

    p = _Parser(_tokens(text))
    if p.peek(1)[1] == "=":
        p.i += 2
    kind, ctor = p.take()
    if ctor not in ("NewOperator", "NewWavefunction"):
        raise QSError(f"io.get reads NewOperator and NewWavefunction texts, not {ctor}")
    p.take("(")
    args = [p.value()]
    while p.peek()[1] == ",":
        p.take(",")
        args.append(p.value())
    p.take(")")
    nf, nb = int(args[0]), int(args[1])
    opts = {o[0]: o[1] for o in (args[3] if len(args) > 3 else []) if isinstance(o, list) and len(o) == 2}
    if ctor == "NewOperator":
        keep = [[k, v] for k, v in opts.items() if k in ("Name", "Restrictions")]
        return NewOperator(nf, nb, args[2], keep or None)
    return NewWavefunction(nf, nb, [(s, c) for s, c in args[2]])


def get(dirname: str, name: str):
    """io.get: reads dirname/name.Quanty written by put (or by the reference)."""
    with open(os.path.join(dirname, name + ".Quanty")) as f:
        return parse(f.read())
