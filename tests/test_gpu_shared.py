"""Bindings that share an assembled connection stream (qs_hop_create_like):
an operator of the same structure with other values, bound to a cached
stream, gives bit for bit the H*psi, diagonal and layout of its own fresh
assembly, agrees with the oracle's push-form H*psi, outlives the binding it
was made from, and is refused when its terms act differently. The server's
pattern cache serves the same spectra as serving without it."""
import gc

import numpy as np
import pytest
import torch

import oracle
from quanty_amd import det as D
from quanty_amd import problems as P
from quanty_amd import serve
from quanty_amd._lib import QSError
from quanty_amd.engine import Basis, Hop, Op

pytestmark = pytest.mark.gpu

SHAPE = dict(nbath=2, full_baths=1, nelec=18, d_range=(6, 10), max_holes=1, max_elec=1)


def _x(n, seed):
    g = np.random.default_rng(seed)
    return torch.from_numpy(g.normal(size=n) + 1j * g.normal(size=n)).cuda()


def test_shared_stream_equals_fresh_assembly_and_the_oracle():
    a, b = P.anderson(seed=3, **SHAPE), P.anderson(seed=4, **SHAPE)
    nf = a["NF"]
    assert [c for c, _ in a["H"]] == [c for c, _ in b["H"]] and a["H"] != b["H"]
    basis = Basis.enumerate(nf, a["restrictions"])
    ha = Hop(Op(nf, a["H"]), basis, mode="assembled")
    opb = Op(nf, b["H"])
    hb = Hop.like(opb, ha)
    fresh = Hop(Op(nf, b["H"]), Basis.enumerate(nf, b["restrictions"]), mode="assembled")
    sf, sl = fresh.stats(), hb.stats()
    assert sf == sl
    assert torch.equal(hb.diagonal(), fresh.diagonal())
    x = _x(basis.n, 1)
    y = hb.apply(x)
    assert torch.equal(y, fresh.apply(x))
    assert not torch.equal(y, ha.apply(x))  # the values are b's, not a's
    # the template may go first: the stream lives while a binding uses it
    del ha
    gc.collect()
    assert torch.equal(hb.apply(x), y)
    # against the oracle's push-form H*psi of b
    dets = D.words_to_bytes(basis.words(), nf)
    xh = x.cpu().numpy()
    d, v = oracle.Op(nf, b["H"]).apply(oracle.Vec.from_arrays(nf, dets, xh), restrictions=b["restrictions"]).arrays()
    index = {bytes(r): i for i, r in enumerate(dets)}
    want = np.zeros(basis.n, dtype=complex)
    for row, val in zip(d, v):
        want[index[bytes(row)]] += val
    got = y.cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))


def test_shared_stream_is_refused_for_other_terms():
    a = P.anderson(seed=3, **SHAPE)
    nf = a["NF"]
    basis = Basis.enumerate(nf, a["restrictions"])
    ha = Hop(Op(nf, a["H"]), basis, mode="assembled")
    with pytest.raises(QSError, match="terms"):
        Hop.like(Op(nf, a["H"][:-1]), ha)
    other = list(a["H"])
    c, v = other[-1]
    other[-1] = (c, 0.0)  # a vanishing coefficient drops the term when compiled
    with pytest.raises(QSError, match="terms"):
        Hop.like(Op(nf, other), ha)
    mf = Hop(Op(nf, a["H"]), basis, mode="matfree")
    with pytest.raises(QSError, match="assembled"):
        Hop.like(Op(nf, a["H"]), mf)


def test_server_pattern_cache_serves_the_same_spectra():
    from quanty_amd import default_context
    ctx = default_context()
    reqs = serve.request_stream(24, 0.0, seed=3, dim_lo=2e3, dim_hi=6e3)
    shapes = {tuple(sorted((k, str(v)) for k, v in r.shape.items())) for r in reqs}
    assert len(shapes) < len(reqs)  # some shapes repeat, so the cache is hit
    plain = serve.SpectrumServer(ctx, max_slots=16, share_patterns=False).serve(
        serve.request_stream(24, 0.0, seed=3, dim_lo=2e3, dim_hi=6e3), realtime=False)
    serve.PATTERNS.clear()
    shared = serve.SpectrumServer(ctx, max_slots=16).serve(reqs, realtime=False)
    # two preparation workers may both miss on one shape before either stores it
    assert serve.PATTERNS.hits + serve.PATTERNS.misses == len(reqs)
    assert serve.PATTERNS.hits >= (len(reqs) - len(shapes)) // 2
    by = {r.rid: r for r in plain}
    for r in shared:
        assert np.array_equal(r.spectrum, by[r.rid].spectrum), r.rid
        assert r.E0 == by[r.rid].E0
