"""Host-side pieces of libqs and the C ABI surface (no GPU needed)."""
import os
import re

import numpy as np
import pytest

from quanty_amd import _lib
from quanty_amd.engine import Basis, heev_host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    hdr = open(os.path.join(ROOT, "include", "qs.h")).read()
    declared = set(re.findall(r"\b(qs_[a-z0-9_]+)\s*\(", hdr))
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(lib, name), f"libqs does not export {name}"
    assert declared == set(_lib.EXPORTED), sorted(declared ^ set(_lib.EXPORTED))


@pytest.mark.parametrize("n", [1, 2, 5, 17, 64, 150])
def test_heev_matches_numpy(n):
    rng = np.random.default_rng(n)
    M = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A = M + M.conj().T
    w, Z = heev_host(A)
    assert np.allclose(w, np.linalg.eigvalsh(A), atol=1e-11 * max(1, np.abs(w).max()))
    assert np.allclose(Z.conj().T @ Z, np.eye(n), atol=1e-11)
    assert np.allclose(A @ Z, Z * w, atol=1e-10 * max(1, np.abs(w).max()))


def test_heev_degenerate_and_tridiagonal():
    n = 40
    d = np.repeat(np.arange(10.0), 4)
    Q, _ = np.linalg.qr(np.random.default_rng(1).normal(size=(n, n)) + 1j * np.random.default_rng(2).normal(size=(n, n)))
    A = (Q * d) @ Q.conj().T
    w, Z = heev_host(A)
    assert np.allclose(w, d, atol=1e-11)
    assert np.allclose(A @ Z, Z * w, atol=1e-10)
    T = np.diag(np.linspace(-1, 1, 30)) + np.diag(np.full(29, 0.3), 1) + np.diag(np.full(29, 0.3), -1)
    w, Z = heev_host(T)
    assert np.allclose(w, np.linalg.eigvalsh(T), atol=1e-13)


def test_rank_table_counts_match_oracle():
    import oracle
    cases = [
        (20, [("1" * 20, 8, 8)]),
        (24, [("1" * 24, 10, 10), ("111111000000111111000000", 2, 5), ("000111111111111000000000", 4, 9)]),
        (30, [("1" * 30, 12, 14), ("101010101010101010101010101010", 5, 9)]),
        (17, []),
    ]
    for nf, restr in cases:
        assert Basis.count(nf, restr) == len(oracle.enumerate_sector(nf, restr))


@pytest.mark.parametrize("nres", [1, 3, 8, 9, 11])
def test_rank_table_random_restrictions(nres):
    """Random overlapping masks: 8 restrictions or fewer take the packed-state
    table builder, more take the general one; both against brute force over
    2^nf determinants (and bounds beyond the mode count)."""
    rng = np.random.default_rng(nres)
    nf = 14
    for trial in range(6):
        restr = []
        for _ in range(nres):
            m = rng.integers(0, 2, nf)
            if m.sum() == 0:
                m[rng.integers(nf)] = 1
            k = int(m.sum())
            lo = int(rng.integers(0, k + 1))
            hi = int(rng.integers(lo, k + 1)) if trial % 3 else 400  # an upper bound past every count
            restr.append(("".join(map(str, m)), lo, hi))
        dets = np.arange(1 << nf)
        bits = (dets[:, None] >> (nf - 1 - np.arange(nf))) & 1
        ok = np.ones(len(dets), dtype=bool)
        for mstr, lo, hi in restr:
            c = bits @ np.array([int(ch) for ch in mstr])
            ok &= (c >= lo) & (c <= hi)
        assert Basis.count(nf, restr) == int(ok.sum()), restr


@pytest.mark.parametrize("m", [1, 2, 7, 60, 400])
def test_tridiagonal_lowest_and_extremes(m):
    from quanty_amd.engine import tridiag_extremes, tridiag_lowest
    rng = np.random.default_rng(m)
    d = rng.normal(size=m) * 3
    e = np.abs(rng.normal(size=m - 1)) + (1e-3 if m > 2 else 0.0)
    if m > 5:
        e[m // 2] = 0.0  # a split block
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w, V = np.linalg.eigh(T)
    E, z = tridiag_lowest(d, e)
    assert abs(E - w[0]) < 1e-12 * max(1, abs(w[0]))
    assert np.linalg.norm(T @ z - E * z) < 1e-9 * max(1, np.abs(w).max())
    assert abs(np.linalg.norm(z) - 1) < 1e-12
    lo, hi = tridiag_extremes(d, e)
    assert abs(lo - w[0]) < 1e-12 * max(1, abs(w[0])) and abs(hi - w[-1]) < 1e-12 * max(1, abs(w[-1]))


def test_tridiagonal_lowest_with_bracket_hints():
    """The bracketed search returns the eigenvalue numpy finds whether the
    hints are right, loose or wrong (wrong ones are detected and ignored),
    including Lanczos tables with ghost copies of the lowest eigenvalue."""
    from quanty_amd.engine import tridiag_lowest
    g = np.random.default_rng(7)
    for trial in range(300):
        m = int(g.integers(2, 160))
        if trial % 2:
            d = g.normal(size=m) * 5 + 50
            e = np.abs(g.normal(size=m - 1)) * 2
            e[g.random(m - 1) < 0.1] = 0.0
        else:  # Lanczos without re-orthogonalization on a small matrix
            n = max(4, m // 4)
            A = g.normal(size=(n, n))
            A = A + A.T
            v = g.normal(size=n)
            v /= np.linalg.norm(v)
            vp, bprev = np.zeros(n), 0.0
            d, e = np.zeros(m), np.zeros(m - 1)
            for k in range(m):
                w = A @ v - bprev * vp
                d[k] = v @ w
                w -= d[k] * v
                if k < m - 1:
                    bprev = max(np.linalg.norm(w), 1e-300)
                    e[k] = bprev
                    vp, v = v, w / bprev
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        w0 = np.linalg.eigvalsh(T)[0]
        s = max(1.0, np.abs(np.linalg.eigvalsh(T)).max())
        for lo, hi in [(None, None), (w0 - 1e-4 * s, w0 + 1e-9 * s), (w0 + 0.3 * s, w0 - 0.3 * s),
                       (w0 - 10 * s, None), (None, w0 + 1e-14 * s)]:
            E, z = tridiag_lowest(d, e, lo, hi)
            assert abs(E - w0) < 1e-13 * s, (trial, lo, hi, E, w0)
            assert np.linalg.norm(T @ z - E * z) < 1e-11 * s


def test_device_memory_arena_bookkeeping(tmp_path):
    """The sub-allocator behind the library's device memory (qs_arena.h),
    compiled for the host: randomized take/give over adjacent chunks keeps live
    blocks disjoint and inside their chunks, merges freed neighbours (never
    across chunk edges) and accounts every byte."""
    import subprocess
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "arena_check.cpp")
    exe = str(tmp_path / "arena_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-o", exe, src], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert "arena ok" in out.stdout
