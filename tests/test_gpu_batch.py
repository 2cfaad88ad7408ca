"""Batched Lanczos and the continuous-batching spectrum server on the GPU."""
import numpy as np
import pytest
import torch

import oracle
from quanty_amd import det as D
from quanty_amd import problems as P
from quanty_amd import serve
from quanty_amd.engine import Basis, Batch, Hop, Op, heev_host

pytestmark = pytest.mark.gpu

SHAPES = [
    dict(seed=3, nbath=1, nelec=12, full_baths=1, max_holes=3, max_elec=0, d_range=(0, 10)),
    dict(seed=5, nbath=2, nelec=18, full_baths=1, max_holes=2, max_elec=2),
    dict(seed=9, nbath=3, nelec=22, full_baths=1, max_holes=2, max_elec=3),
    dict(seed=11, nbath=2, nelec=17, full_baths=1, max_holes=1, max_elec=1, d_range=(5, 9)),
]


def _setup(kw):
    pr = P.anderson(**kw)
    b = Basis.enumerate(pr["NF"], pr["restrictions"])
    hop = Hop(Op(pr["NF"], pr["H"]), b, mode="assembled")
    g = np.random.default_rng(kw["seed"])
    v0 = torch.from_numpy(g.normal(size=b.n) + 1j * g.normal(size=b.n)).cuda()
    return pr, b, hop, v0


def test_batch_matches_single_and_is_mix_invariant():
    probs = [_setup(kw) for kw in SHAPES]
    depths = [40, 55, 30, 47]
    mixed = Batch(8)
    slots = [mixed.add(hop, v0, d) for (_, _, hop, v0), d in zip(probs, depths)]
    for _ in range(8):
        mixed.run(7)
    for (pr, b, hop, v0), d, s in zip(probs, depths, slots):
        st, a, bb = mixed.poll(s)
        assert st == d
        # alone: bit-identical coefficients
        alone = Batch(1)
        s1 = alone.add(hop, v0, d)
        alone.run(d + 5)
        st1, a1, b1 = alone.poll(s1)
        assert np.array_equal(a, a1) and np.array_equal(bb, b1)
        # against the single-problem Lanczos and the oracle (epsilon = 0)
        ra, rb = hop.lanczos(v0, d)
        assert np.allclose(a, ra, rtol=1e-10, atol=1e-12)
        assert np.allclose(bb[:d], rb, rtol=1e-10, atol=0)
        if b.n > 20000:
            continue  # the oracle comparison is kept to the smaller problems
        x = v0.cpu().numpy()
        psi = oracle.Vec.from_arrays(pr["NF"], D.words_to_bytes(b.words(), pr["NF"]), x)
        oa, ob = oracle.Op(pr["NF"], pr["H"]).lanczos(psi, d, restrictions=pr["restrictions"])
        assert np.allclose(a, oa, rtol=1e-10, atol=1e-12) and np.allclose(bb[:d], ob, rtol=1e-10)


def test_batch_mixed_layouts():
    """Slots with unpacked entries and with two-word (76-mode) determinants
    next to ordinary slots: every slot matches its single-problem Lanczos and
    the ordinary slots are bit-identical to a run without the others."""
    from quanty_amd import default_context
    from quanty_amd._lib import QS_PARAM_ELL_PACK, QS_PARAM_SHIFT_MIN

    ctx = default_context()
    plain = [_setup(kw) for kw in SHAPES[:2]]
    ctx.set_param(QS_PARAM_ELL_PACK, 0)
    ctx.set_param(QS_PARAM_SHIFT_MIN, 8)
    try:
        unpacked = _setup(SHAPES[3])
    finally:
        ctx.set_param(QS_PARAM_ELL_PACK, 1)
        ctx.set_param(QS_PARAM_SHIFT_MIN, 56)
    assert not unpacked[2].stats()["packed"]
    # 76 modes: two-word determinants, a p-d hopping operator on a small sector
    nf = 76
    restr = [("1" * nf, 3, 3), ("1" * 20 + "0" * 56, 1, 2), ("0" * 10 + "1" * 60 + "0" * 6, 1, 2)]
    b76 = Basis.enumerate(nf, restr)
    assert b76.nw == 2 and b76.n == 26520
    rng = np.random.default_rng(3)
    terms = []
    for _ in range(60):
        i, j = (int(q) for q in rng.integers(0, nf, 2))
        if i != j:
            v = complex(*rng.normal(size=2))
            terms += [((0x8000 | i, j)[::-1], v), ((0x8000 | j, i)[::-1], v.conjugate())]
    h76 = Hop(Op(nf, terms), b76, mode="assembled")
    g = np.random.default_rng(76)
    v76 = torch.from_numpy(g.normal(size=b76.n) + 1j * g.normal(size=b76.n)).cuda()
    probs = plain + [unpacked, (None, b76, h76, v76)]
    depths = [30, 25, 35, 20]
    mixed = Batch(8)
    slots = [mixed.add(hop, v0, d) for (_, _, hop, v0), d in zip(probs, depths)]
    mixed.run(40)
    for k, ((_, b, hop, v0), d, s) in enumerate(zip(probs, depths, slots)):
        st, a, bb = mixed.poll(s)
        assert st == d
        ra, rb = hop.lanczos(v0, d)
        assert np.allclose(a, ra, rtol=1e-10, atol=1e-12)
        assert np.allclose(bb[:d], rb, rtol=1e-10, atol=0)
        if k < 2:
            alone = Batch(2)
            s1 = alone.add(hop, v0, d)
            alone.run(d)
            _, a1, b1 = alone.poll(s1)
            assert np.array_equal(a, a1) and np.array_equal(bb, b1)


def test_two_pass_ritz_vector():
    """Coefficients, then the Ritz vector rebuilt from the same start: the
    lowest eigenpair of the tridiagonal matrix becomes an eigenvector of H."""
    pr, b, hop, v0 = _setup(SHAPES[0])
    m = 300
    bt = Batch(3)
    s = bt.add(hop, v0, m)
    bt.run(m)
    _, a, bb = bt.poll(s)
    T = np.diag(a[1:m + 1]).astype(complex)
    for i in range(1, m):
        T[i, i - 1] = T[i - 1, i] = bb[i]
    w, Z = heev_host(T)
    z = (Z[:, 0] * np.exp(-1j * np.angle(Z[np.argmax(np.abs(Z[:, 0])), 0]))).real
    s2 = bt.add(hop, v0, m, ycoef=z)
    bt.run(m)
    y = b.zeros()
    bt.poll(s2, want_coef=False, y=y)
    # the same Ritz vector from the kept Lanczos vectors of a one-pass slot
    s3 = bt.add(hop, v0, m, keep_basis=True)
    bt.run(m)
    _, a3, b3 = bt.poll(s3)
    assert np.array_equal(a3, a) and np.array_equal(b3, bb)
    yk = bt.ritz(s3, z, b.zeros())
    assert torch.equal(yk, y)
    y /= torch.linalg.vector_norm(y)
    hy = hop.apply(y.contiguous())
    res = torch.linalg.vector_norm(hy - w[0] * y).item()
    assert res < 1e-6
    E, _, _, _ = hop.ground_states(1, tol=1e-14)
    assert abs(E[0] - w[0]) < 1e-9 * abs(E[0])


def test_server_results_same_alone_and_in_the_mix():
    """A stream of requests of different sizes: every spectrum produced in the
    continuous batch equals the one produced by the plain one-at-a-time run, bit
    for bit."""
    from quanty_amd import default_context
    reqs = serve.request_stream(10, rate=0.0, seed=3, dim_lo=1e2, dim_hi=4e4)
    srv = serve.SpectrumServer(default_context(), max_slots=16, gs_depth=40, chunk=5)
    mixed = {r.rid: r.spectrum for r in srv.serve(reqs, realtime=False)}
    reqs2 = serve.request_stream(10, rate=0.0, seed=3, dim_lo=1e2, dim_hi=4e4)
    plain = {r.rid: r.spectrum for r in serve.SpectrumServer(default_context(), max_slots=1, gs_depth=40,
                                                             chunk=5).serve(reqs2, realtime=False, plain=True)}
    assert len(mixed) == len(plain) == 10
    for rid in mixed:
        assert np.array_equal(mixed[rid], plain[rid]), rid
        assert np.all(np.isfinite(mixed[rid]))


def test_server_spectrum_matches_oracle():
    """One served request re-done by the oracle from the served ground state."""
    from quanty_amd import default_context
    reqs = serve.request_stream(1, rate=0.0, seed=11, dim_lo=800, dim_hi=1500)
    srv = serve.SpectrumServer(default_context(), max_slots=4, chunk=8)
    keep = {}
    orig = srv._advance

    def spy(job, *args):
        out = orig(job, *args)
        if job.phase == "spec" and "psi" not in keep:
            keep["basis"] = job.basis.words()
            keep["psi"] = job.gs_vector.cpu().numpy()
            keep["nf"] = job.nf
            keep["req"] = job.req
        return out

    srv._advance = spy
    (r,) = srv.serve(reqs, realtime=False)
    pr = P.anderson(seed=r.seed, **r.shape)
    nf = keep["nf"]
    psi = oracle.Vec.from_arrays(nf, D.words_to_bytes(keep["basis"], nf), keep["psi"])
    rng = P.SplitMix64(r.seed ^ 0x5DEECE66D)
    bath = [m for b in pr["baths"] for m in b]
    T = P.random_one_body_transition(rng, pr["dmodes"], bath)
    tpsi = oracle.Op(nf, T).apply(psi, restrictions=pr["restrictions"])
    a, b = oracle.Op(nf, pr["H"]).lanczos(tpsi, r.depth, restrictions=pr["restrictions"])
    want = oracle.cf(a, b, r.E0, srv.emin, srv.emax, srv.ne, r.gamma)
    assert np.max(np.abs(r.spectrum - want)) <= 1e-8 * np.max(np.abs(want))
    assert r.gs_residual ** 2 <= 1e-12 * max(1.0, r.E0 ** 2)


def test_server_kept_basis_equals_second_pass():
    """The ground state formed from the kept Lanczos vectors gives the same
    spectra as the two-pass rebuild."""
    from quanty_amd import default_context
    reqs = serve.request_stream(6, rate=0.0, seed=5, dim_lo=1e2, dim_hi=2e4)
    kept = {r.rid: r for r in serve.SpectrumServer(default_context(), max_slots=8, gs_depth=40,
                                                   chunk=5).serve(reqs, realtime=False)}
    reqs2 = serve.request_stream(6, rate=0.0, seed=5, dim_lo=1e2, dim_hi=2e4)
    two = {r.rid: r for r in serve.SpectrumServer(default_context(), max_slots=8, gs_depth=40, chunk=5,
                                                  keep_basis=False).serve(reqs2, realtime=False)}
    assert len(kept) == len(two) == 6
    for rid, r in kept.items():
        assert r.E0 == two[rid].E0 and r.gs_steps == two[rid].gs_steps
        assert np.array_equal(r.spectrum, two[rid].spectrum), rid



def _graph_sequence(graph: int):
    """Slots joining and leaving between runs of 6 steps, with other device
    work in between; returns every finished slot's coefficients."""
    from quanty_amd import default_context
    from quanty_amd._lib import QS_PARAM_BATCH_GRAPH

    ctx = default_context()
    old = ctx.get_param(QS_PARAM_BATCH_GRAPH)
    ctx.set_param(QS_PARAM_BATCH_GRAPH, graph)
    try:
        probs = [_setup(kw) for kw in SHAPES]
        B = Batch(4, ctx)
        out = {}
        live = {B.add(probs[0][2], probs[0][3], 30): 0, B.add(probs[1][2], probs[1][3], 24): 1}
        joins = [(2, 36), (3, 18), (0, 12)]
        for rnd in range(12):
            B.run(6)
            # unrelated device work between replays
            y = probs[1][2].apply(probs[1][3])
            torch.cuda.synchronize()
            assert torch.isfinite(y.real).all()
            for s in list(live):
                st, a, b = B.poll(s)
                if st >= B.maxdepth[s]:
                    out[(rnd, live[s])] = (a.copy(), b.copy())
                    B.remove(s)
                    del live[s]
            if joins and len(live) < 3:
                i, d = joins.pop(0)
                live[B.add(probs[i][2], probs[i][3], d)] = i
        assert not live and not joins
        return out
    finally:
        ctx.set_param(QS_PARAM_BATCH_GRAPH, old)


def test_batch_graph_replays_match_eager():
    """Graph launches (replayed across slot changes) give the coefficients of
    kernel-by-kernel launches, bit for bit."""
    eager = _graph_sequence(0)
    graph = _graph_sequence(1)
    assert eager.keys() == graph.keys() and len(eager) == 5
    for k in eager:
        assert np.array_equal(eager[k][0], graph[k][0]) and np.array_equal(eager[k][1], graph[k][1]), k


def test_server_weighted_start_reaches_the_same_ground_state():
    """The ground-state start weighted towards low-energy determinants (with a
    floor that keeps every determinant) converges to the energy of the plain
    random start: both meet |H psi - E psi|^2 <= 1e-12 max(1, E^2), so their
    Ritz values differ by at most residual^2 / gap, here below 1e-8 relative."""
    from quanty_amd import default_context
    ctx = default_context()
    reqs = serve.request_stream(6, rate=0.0, seed=9, dim_lo=1e3, dim_hi=3e4)
    a = {r.rid: r for r in serve.SpectrumServer(ctx, max_slots=8, start_temperature=None).serve(reqs, realtime=False)}
    reqs2 = serve.request_stream(6, rate=0.0, seed=9, dim_lo=1e3, dim_hi=3e4)
    b = {r.rid: r for r in serve.SpectrumServer(ctx, max_slots=8).serve(reqs2, realtime=False)}
    assert a.keys() == b.keys()
    for rid in a:
        for r in (a[rid], b[rid]):
            assert r.gs_residual ** 2 <= 1e-12 * max(1.0, r.E0 ** 2)
        assert abs(a[rid].E0 - b[rid].E0) <= 1e-8 * max(1.0, abs(a[rid].E0)), (rid, a[rid].E0, b[rid].E0)


def test_server_restarts_ground_state_when_the_kept_basis_is_full():
    """With room for only 40 kept Lanczos vectors, the ground state is restarted
    from the Ritz vector until it meets the criterion; the spectrum then agrees
    with the oracle's from the served ground state, and the plain run is
    identical."""
    from quanty_amd import default_context
    reqs = serve.request_stream(1, rate=0.0, seed=11, dim_lo=800, dim_hi=1500)
    srv = serve.SpectrumServer(default_context(), max_slots=4, chunk=8, gs_max=40)
    keep = {}
    orig = srv._advance

    def spy(job, *args):
        out = orig(job, *args)
        if job.phase == "spec" and "psi" not in keep:
            keep["basis"] = job.basis.words()
            keep["psi"] = job.gs_vector.cpu().numpy()
            keep["nf"] = job.nf
            keep["restarts"] = job.restarts
        return out

    srv._advance = spy
    (r,) = srv.serve(reqs, realtime=False)
    assert keep["restarts"] >= 1 and r.gs_steps > 40
    assert r.gs_residual ** 2 <= 1e-12 * max(1.0, r.E0 ** 2)
    pr = P.anderson(seed=r.seed, **r.shape)
    nf = keep["nf"]
    psi = oracle.Vec.from_arrays(nf, D.words_to_bytes(keep["basis"], nf), keep["psi"])
    rng = P.SplitMix64(r.seed ^ 0x5DEECE66D)
    bath = [m for b in pr["baths"] for m in b]
    T = P.random_one_body_transition(rng, pr["dmodes"], bath)
    tpsi = oracle.Op(nf, T).apply(psi, restrictions=pr["restrictions"])
    a, b = oracle.Op(nf, pr["H"]).lanczos(tpsi, r.depth, restrictions=pr["restrictions"])
    want = oracle.cf(a, b, r.E0, srv.emin, srv.emax, srv.ne, r.gamma)
    assert np.max(np.abs(r.spectrum - want)) <= 1e-8 * np.max(np.abs(want))
    plain = serve.SpectrumServer(default_context(), max_slots=1, chunk=8, gs_max=40)
    (p,) = plain.serve(serve.request_stream(1, rate=0.0, seed=11, dim_lo=800, dim_hi=1500), realtime=False,
                       plain=True)
    assert np.array_equal(p.spectrum, r.spectrum) and p.gs_steps == r.gs_steps
