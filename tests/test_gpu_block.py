"""Block Lanczos, tensor continued fraction and lowest eigenstates on the GPU,
against dense linear algebra on the oracle's Hamiltonian matrix and the
reference's golden energies."""
import numpy as np
import pytest
import torch

import oracle
from models import f_shell, golden
from quanty_amd import det as D
from quanty_amd import problems as P
from quanty_amd.engine import Basis, Hop, Op, block_cf_spectrum, cf_spectrum, dot

pytestmark = pytest.mark.gpu


def dense_matrix(nf, terms, basis: Basis, restrictions):
    """<d_i|H|d_j> from the oracle's push-form H applied to every basis determinant."""
    words = basis.words()
    byts = D.words_to_bytes(words, nf)
    op = oracle.Op(nf, terms)
    n = basis.n
    H = np.zeros((n, n), dtype=complex)
    for j in range(n):
        v = oracle.Vec(nf)
        v.add(byts[j], 1.0)
        d, val = op.apply(v, restrictions=restrictions).arrays()
        if len(val):
            H[basis.index(D.bytes_to_words(d, nf)), j] += val
    return H


SMALL_GS = dict(seed=21, nbath=1, nelec=17, full_baths=1, max_holes=2, max_elec=0, d_range=(0, 10))


def test_ground_states_match_dense():
    pr = P.anderson(**SMALL_GS)
    nf = pr["NF"]
    b = Basis.enumerate(nf, pr["restrictions"])
    assert b.n == 1020
    H = dense_matrix(nf, pr["H"], b, pr["restrictions"])
    assert np.allclose(H, H.conj().T, atol=1e-12)
    want = np.linalg.eigvalsh(H)[:3]
    for mode, mk in [("assembled", 60), ("matfree", 500)]:
        hop = Hop(Op(nf, pr["H"]), b, mode=mode)
        E, vecs, var, it = hop.ground_states(3, tol=1e-14, max_krylov=mk)
        assert np.allclose(E, want, rtol=1e-10, atol=1e-10), (mode, E, want)
        V = vecs.cpu().numpy()
        bound = 1e-14 * max(1.0, float(np.sum(E ** 2))) / 3
        for i in range(3):
            r = H @ V[i] - E[i] * V[i]
            # the returned variance is the true |H v - E v|^2 and meets the criterion
            assert np.linalg.norm(r) ** 2 <= bound * 1.01
            assert abs(np.linalg.norm(r) ** 2 - var[i]) <= 1e-3 * bound
            assert abs(np.vdot(V[i], V[i]) - 1) < 1e-10
        assert np.allclose(V.conj() @ V.T, np.eye(3), atol=1e-10)


@pytest.mark.parametrize("cplx", [False, True])
def test_ground_states_f6_mathematica(cplx):
    """The 10 lowest f6 states (dim 3003) against the Mathematica energies of
    TestScripts/Eigensystem.Quanty:571 (1e-10 relative) and :669 (reference tolerances)."""
    nf, terms = f_shell(cplx)
    b = Basis.enumerate(nf, [("1" * nf, 6, 6)])
    assert b.n == 3003
    hop = Hop(Op(nf, terms), b, mode="assembled")
    E, vecs, var, it = hop.ground_states(10, tol=1e-15, max_krylov=400)
    want = np.array(golden("eigensystem_f6.json")["complex" if cplx else "real"])
    if not cplx:
        assert np.allclose(E, want, rtol=1e-10, atol=0), (E - want)
    else:
        assert np.all(np.abs(E - want) < 1e-4) and abs(E[0] - want[0]) < 1e-6


def test_ground_states_ni_xas_exhausted_krylov():
    """190-dim Ni2+ ground sector: the Krylov space is exhausted, Ritz values are exact."""
    ni = P.ni_xas()
    b = Basis.enumerate(26, ni["gs_restrictions"])
    assert b.n == 190
    H = dense_matrix(26, ni["Hi"], b, ni["gs_restrictions"])
    want = np.linalg.eigvalsh(H)[:3]
    hop = Hop(Op(26, ni["Hi"]), b, mode="assembled")
    E, vecs, var, it = hop.ground_states(3, tol=1e-14)
    assert np.allclose(E, want, rtol=1e-11, atol=1e-11)


def test_block_lanczos_tensor_resolvent():
    """With N*k >= dim the block Krylov space is exhausted and the tensor CF equals
    X^+ (z - H)^-1 X exactly; its diagonal equals the scalar Lanczos CF."""
    ni = P.ni_xas()
    fs = Basis.enumerate(26, ni["fs_restrictions"])  # 120 dims
    H = dense_matrix(26, ni["Hf"], fs, ni["fs_restrictions"])
    hop = Hop(Op(26, ni["Hf"]), fs, mode="assembled")
    rng = np.random.default_rng(4)
    k = 3
    X = rng.normal(size=(k, fs.n)) + 1j * rng.normal(size=(k, fs.n))
    A, B, kept = hop.block_lanczos(torch.from_numpy(X).cuda(), 60)
    assert kept[1] == k
    E0, emin, emax, ne, gamma = 0.0, -20.0, 40.0, 300, 0.6
    G = block_cf_spectrum(A, B, kept, E0, emin, emax, ne, gamma)
    w = emin + np.arange(ne + 1) * (emax - emin) / ne
    for i in range(0, ne + 1, 37):
        z = w[i] + E0 + 0.5j * gamma
        R = X.conj() @ np.linalg.solve(z * np.eye(fs.n) - H, X.T)
        assert np.max(np.abs(G[i] - R)) < 1e-9 * np.max(np.abs(R))
    for r in range(k):
        a, bb = hop.lanczos(torch.from_numpy(X[r].copy()).cuda(), 120)
        g = cf_spectrum(a, bb, E0, emin, emax, ne, gamma)[0]
        assert np.max(np.abs(g - G[:, r, r])) < 1e-8 * np.max(np.abs(g))


def test_block_lanczos_rank_deficient_start():
    """Linearly dependent start vectors are reduced (Loewdin cut) and the
    tensor spectrum still equals the dense resolvent."""
    ni = P.ni_xas()
    fs = Basis.enumerate(26, ni["fs_restrictions"])
    H = dense_matrix(26, ni["Hf"], fs, ni["fs_restrictions"])
    hop = Hop(Op(26, ni["Hf"]), fs, mode="matfree")
    rng = np.random.default_rng(5)
    x = rng.normal(size=(2, fs.n)) + 1j * rng.normal(size=(2, fs.n))
    X = np.stack([x[0], x[1], x[0] + 2j * x[1]])
    A, B, kept = hop.block_lanczos(torch.from_numpy(X).cuda(), 80)
    assert kept[1] == 2
    G = block_cf_spectrum(A, B, kept, 0.0, -20.0, 40.0, 60, 0.8)
    z = -20.0 + 0.4j
    R = X.conj() @ np.linalg.solve(z * np.eye(fs.n) - H, X.T)
    assert np.max(np.abs(G[0] - R)) < 1e-9 * np.max(np.abs(R))


def test_ground_states_full_size_anderson():
    """The 2,179,235-dim workload: 3 lowest states to a 1e-12 variance criterion."""
    pr = P.anderson(seed=1)
    nf = pr["NF"]
    b = Basis.enumerate(nf, pr["restrictions"])
    hop = Hop(Op(nf, pr["H"]), b, mode="assembled")
    E, vecs, var, it = hop.ground_states(3, tol=1e-12, max_krylov=300)
    assert np.all(np.diff(E) >= -1e-12)
    assert np.all(var <= 1e-12 * max(1.0, float(np.sum(E ** 2))) / 3 * 1.0001)
    hv = hop.apply(vecs[0].contiguous())
    e0 = dot(vecs[0].contiguous(), hv).real
    assert abs(e0 - E[0]) < 1e-9 * abs(E[0])


def test_ground_states_krylov_schedules_agree():
    """The Krylov growth schedule (start, step, maximum basis) changes how the
    search proceeds, not its answer: every schedule reaches the dense
    eigenvalues, including a fixed 40-vector basis that restarts every few
    blocks."""
    pr = P.anderson(**SMALL_GS)
    nf = pr["NF"]
    b = Basis.enumerate(nf, pr["restrictions"])
    want = np.linalg.eigvalsh(dense_matrix(nf, pr["H"], b, pr["restrictions"]))[:3]
    hop = Hop(Op(nf, pr["H"]), b, mode="assembled")
    for start, step, mk in [(100, 10, 500), (20, 5, 60), (37, 0, 40), (300, 10, 500)]:
        E, vecs, var, it = hop.ground_states(3, tol=1e-14, max_krylov=mk, krylov_start=start, krylov_step=step,
                                             max_restarts=400)
        assert np.allclose(E, want, rtol=1e-10, atol=1e-10), (start, step, mk, E, want)
        for i in range(3):
            v = vecs[i].contiguous()
            assert abs(dot(v, v).real - 1.0) < 1e-10
            r = hop.apply(v) - E[i] * v
            assert dot(r, r).real < 1e-10 * max(1.0, float(np.sum(E ** 2)))


def test_device_memory_is_reused_and_trimmed():
    """Freed library memory stays in the process's chunks and is reused (no
    growth over repeated create/destroy cycles); a block larger than a chunk
    gets a chunk of its own, which qs_trim_memory hands back to the driver
    once it is wholly free; the freed space of a large block serves smaller
    blocks."""
    import ctypes as C
    from quanty_amd import default_context
    ctx = default_context()
    pr = P.anderson(**SMALL_GS)
    nf = pr["NF"]
    b = Basis.enumerate(nf, pr["restrictions"])
    op = Op(nf, pr["H"])
    warm = Hop(op, b, mode="assembled")  # first use may grow the context's scratch space
    del warm
    live0, free0 = ctx.memory_stats()
    hop = Hop(op, b, mode="assembled")
    live1, _ = ctx.memory_stats()
    assert live1 > live0
    del hop
    live2, free2 = ctx.memory_stats()
    assert live2 == live0 and free2 == free0
    for _ in range(3):  # the same space comes back
        hop = Hop(op, b, mode="assembled")
        del hop
        assert ctx.memory_stats() == (live0, free0)
    big = free0 + (6 << 30)  # more than the free space held: a chunk of its own
    p = C.c_void_p()
    ctx.check(ctx.lib.qs_dev_alloc(ctx.h, big, C.byref(p)), "qs_dev_alloc")
    assert ctx.memory_stats()[0] >= live0 + big
    ctx.check(ctx.lib.qs_dev_free(ctx.h, p), "qs_dev_free")
    live, free = ctx.memory_stats()
    assert live == live0 and free >= free0 + big
    # two blocks of 2 GiB fit into the freed chunk without a new one
    q1, q2 = C.c_void_p(), C.c_void_p()
    ctx.check(ctx.lib.qs_dev_alloc(ctx.h, 2 << 30, C.byref(q1)), "qs_dev_alloc")
    ctx.check(ctx.lib.qs_dev_alloc(ctx.h, 2 << 30, C.byref(q2)), "qs_dev_alloc")
    assert ctx.memory_stats()[1] == free - (4 << 30)
    ctx.check(ctx.lib.qs_dev_free(ctx.h, q1), "qs_dev_free")
    ctx.check(ctx.lib.qs_dev_free(ctx.h, q2), "qs_dev_free")
    assert ctx.memory_stats() == (live, free)
    ctx.trim_memory(0)
    live3, free3 = ctx.memory_stats()
    assert live3 == live0 and free3 <= free - big
    hop = Hop(op, b, mode="assembled")  # still works after the trim
    assert ctx.memory_stats()[0] == live1
    del hop
