"""Product-layout bindings (qs_hop_create_product) against the oracle and the
SELL layout: H*psi to 1e-13 on every split the restrictions allow (one and
two words per determinant), the same off-diagonal entry count as SELL, Lanczos
coefficients to 1e-10, the order conversions, the diagonal, bindings sharing
an index (create_like) and the refusal of splits the sector does not factor
over, and product slots in the batch engine. All calls go through the C ABI."""
import numpy as np
import pytest
import torch

import oracle
from quanty_amd import det as D
from quanty_amd import problems as P
from quanty_amd._lib import QSError
from quanty_amd.engine import Basis, Hop, Op

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def cvec(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def oracle_apply(nf, terms, basis, x, restrictions):
    psi = oracle.Vec.from_arrays(nf, D.words_to_bytes(basis.words(), nf), x)
    d, v = oracle.Op(nf, terms).apply(psi, restrictions=restrictions).arrays()
    out = np.zeros(basis.n, dtype=complex)
    if len(v):
        idx = basis.index(D.bytes_to_words(d, nf))
        assert np.all(idx >= 0)
        out[idx] = v
    return out


# (problem, group A) pairs: the d shell, a group of bath shells, and a
# two-word case
CASES = [
    (dict(seed=5, nbath=2, nelec=18, full_baths=1, max_holes=2, max_elec=2), "d"),
    (dict(seed=5, nbath=2, nelec=18, full_baths=1, max_holes=2, max_elec=2), "full"),
    (dict(seed=9, nbath=3, nelec=22, full_baths=1, max_holes=2, max_elec=3), "d"),
    (dict(seed=9, nbath=3, nelec=22, full_baths=1, max_holes=2, max_elec=3), "empty"),
    (dict(seed=4, nbath=6, nelec=38, full_baths=3, max_holes=1, max_elec=1), "d"),
]


def group(pr, which, kw):
    """Group A: the d shell, the nominally full bath shells or the empty ones."""
    if which == "d":
        return pr["dmodes"]
    nfull = kw["full_baths"]
    sel = pr["baths"][:nfull] if which == "full" else pr["baths"][nfull:]
    return [m for b in sel for m in b]


@pytest.mark.parametrize("kw,which", CASES)
def test_product_apply_matches_oracle_and_sell(kw, which):
    pr = P.anderson(**kw)
    nf = pr["NF"]
    basis = Basis.enumerate(nf, pr["restrictions"])
    H = Op(nf, pr["H"])
    prod = Hop.product(H, basis, group(pr, which, kw))
    sell = Hop(H, basis, mode="assembled")
    sp, ss = prod.stats(), sell.stats()
    assert sp["assembled"]
    assert sp["nnz_offdiag"] == ss["nnz_offdiag"]
    assert sp["nnz_eff"] == ss["nnz_eff"]
    x = cvec(basis.n, 11)
    want = oracle_apply(nf, pr["H"], basis, x, pr["restrictions"])
    y = prod.apply(torch.from_numpy(x).to(DEV)).cpu().numpy()
    err = np.max(np.abs(y - want)) / np.max(np.abs(want))
    assert err < 1e-13, err
    ys = sell.apply(torch.from_numpy(x).to(DEV)).cpu().numpy()
    assert np.max(np.abs(y - ys)) / np.max(np.abs(ys)) < 1e-13
    # several vectors at once (launches of 4, 2 and 1)
    X = np.stack([cvec(basis.n, 20 + k) for k in range(7)])
    Y = torch.zeros((7, basis.n), dtype=torch.complex128, device=DEV)
    prod.ctx.check(prod.lib.qs_hop_apply_multi(prod.ctx.h, prod.h, 7, torch.from_numpy(X).to(DEV).data_ptr(),
                                               Y.data_ptr()), "qs_hop_apply_multi")
    Ys = sell.apply(torch.from_numpy(X[5]).to(DEV)).cpu().numpy()
    assert np.max(np.abs(Y[5].cpu().numpy() - Ys)) / np.max(np.abs(Ys)) < 1e-13


@pytest.mark.parametrize("kw,which", CASES[:3])
def test_product_lanczos_and_orders(kw, which):
    pr = P.anderson(**kw)
    nf = pr["NF"]
    basis = Basis.enumerate(nf, pr["restrictions"])
    H = Op(nf, pr["H"])
    prod = Hop.product(H, basis, group(pr, which, kw))
    sell = Hop(H, basis, mode="assembled")
    x = cvec(basis.n, 3)
    a1, b1 = sell.lanczos(torch.from_numpy(x).to(DEV), 60)
    a2, b2 = prod.lanczos(torch.from_numpy(x).to(DEV), 60)
    assert np.allclose(a2, a1, rtol=1e-10, atol=1e-10 * np.max(np.abs(a1)))
    assert np.allclose(b2, b1, rtol=1e-10, atol=0)
    psi = oracle.Vec.from_arrays(nf, D.words_to_bytes(basis.words(), nf), x)
    ra, rb = oracle.Op(nf, pr["H"]).lanczos(psi, 30, restrictions=pr["restrictions"])
    assert np.allclose(a2[:31], ra, rtol=1e-10, atol=1e-10) and np.allclose(b2[:30], rb, rtol=1e-10, atol=0)
    # the row order is a permutation; apply_rows equals apply in that order
    order = prod.row_order()
    assert np.array_equal(np.sort(order), np.arange(basis.n))
    xr = torch.from_numpy(np.ascontiguousarray(x[order])).to(DEV)
    yr = prod.apply_rows(xr).cpu().numpy()
    y = sell.apply(torch.from_numpy(x).to(DEV)).cpu().numpy()
    assert np.max(np.abs(yr - y[order])) / np.max(np.abs(y)) < 1e-13
    # the diagonal in canonical order
    assert np.allclose(prod.diagonal().cpu().numpy(), sell.diagonal().cpu().numpy(), rtol=0, atol=1e-12)


def test_product_create_like_binds_new_values():
    kw = dict(seed=5, nbath=2, nelec=18, full_baths=1, max_holes=2, max_elec=2)
    pr, pr2 = P.anderson(**kw), P.anderson(**dict(kw, seed=6))
    nf = pr["NF"]
    basis = Basis.enumerate(nf, pr["restrictions"])
    H1, H2 = Op(nf, pr["H"]), Op(nf, pr2["H"])
    p1 = Hop.product(H1, basis, pr["dmodes"])
    p2 = Hop.like(H2, p1)
    x = cvec(basis.n, 8)
    want = oracle_apply(nf, pr2["H"], basis, x, pr2["restrictions"])
    y = p2.apply(torch.from_numpy(x).to(DEV)).cpu().numpy()
    assert np.max(np.abs(y - want)) / np.max(np.abs(want)) < 1e-13
    del p1  # the shared index outlives the binding it came from
    y = p2.apply(torch.from_numpy(x).to(DEV)).cpu().numpy()
    assert np.max(np.abs(y - want)) / np.max(np.abs(want)) < 1e-13


def test_product_refuses_a_split_that_does_not_factor():
    kw = dict(seed=5, nbath=2, nelec=18, full_baths=1, max_holes=2, max_elec=2)
    pr = P.anderson(**kw)
    nf = pr["NF"]
    basis = Basis.enumerate(nf, pr["restrictions"])
    H = Op(nf, pr["H"])
    # half the d shell: the d-occupation restriction then spans both groups
    with pytest.raises(QSError, match="both groups"):
        Hop.product(H, basis, pr["dmodes"][:5])


def test_product_chooses_a_split_by_its_load_estimate():
    kw = dict(seed=9, nbath=3, nelec=22, full_baths=1, max_holes=2, max_elec=3)
    pr = P.anderson(**kw)
    nf = pr["NF"]
    basis = Basis.enumerate(nf, pr["restrictions"])
    H = Op(nf, pr["H"])
    auto = Hop.product(H, basis)
    costs = {}
    for which in ("d", "full", "empty"):
        costs[which] = Hop.product(H, basis, group(pr, which, kw)).product_info()["load_cost"]
    assert auto.product_info()["load_cost"] == min(costs.values())
    x = cvec(basis.n, 4)
    want = oracle_apply(nf, pr["H"], basis, x, pr["restrictions"])
    y = auto.apply(torch.from_numpy(x).to(DEV)).cpu().numpy()
    assert np.max(np.abs(y - want)) / np.max(np.abs(want)) < 1e-13


def test_product_binding_through_block_krylov_and_ground_state_solvers():
    """The multi-vector entry points reach H*psi through the canonical-order
    conversions of a product binding: their results equal the SELL binding's."""
    kw = dict(seed=9, nbath=3, nelec=22, full_baths=1, max_holes=2, max_elec=3)
    pr = P.anderson(**kw)
    nf = pr["NF"]
    basis = Basis.enumerate(nf, pr["restrictions"])
    H = Op(nf, pr["H"])
    prod = Hop.product(H, basis)
    sell = Hop(H, basis, mode="assembled")
    E1, V1, var1, _ = sell.ground_states(3, tol=1e-12)
    E2, V2, var2, _ = prod.ground_states(3, tol=1e-12)
    assert np.allclose(E2, E1, rtol=1e-10, atol=0)
    # the same states up to a phase: |<v1|v2>| = 1
    ov = torch.abs(torch.sum(V1.conj() * V2, dim=1)).cpu().numpy()
    assert np.all(np.abs(ov - 1.0) < 1e-8), ov
    X = torch.from_numpy(np.stack([cvec(basis.n, 30 + k) for k in range(3)])).to(DEV)
    A1, B1, k1 = sell.block_lanczos(X, 12)
    A2, B2, k2 = prod.block_lanczos(X, 12)
    assert np.array_equal(k1, k2)
    scale = np.max(np.abs(A1))
    assert np.max(np.abs(A2[:8] - A1[:8])) < 1e-10 * scale and np.max(np.abs(B2[:8] - B1[:8])) < 1e-10 * scale
    x = X[0].contiguous()
    a1, b1, W1 = sell.lanczos_krylov(x, 30)
    a2, b2, W2 = prod.lanczos_krylov(x, 30)
    assert np.allclose(a2, a1, rtol=1e-10, atol=1e-10 * np.max(np.abs(a1)))
    assert np.allclose(b2, b1, rtol=1e-10, atol=0)
    assert torch.max(torch.abs(W2 - W1)).item() < 1e-8


def test_product_slots_in_the_batch():
    """Product-layout bindings as batch slots (k_b1p), next to sliced-ELL slots:
    every product slot's coefficients equal its SELL binding's single-problem
    Lanczos to 1e-10 (its sums run in the layout's row order) and are
    bit-identical alone and in the mix; the SELL slots are bit-identical to a
    batch without product slots; the Ritz vector from kept vectors and the
    second-pass vector come back in canonical order and equal the SELL slot's
    to 1e-12."""
    from quanty_amd.engine import Batch, heev_host
    kws = [CASES[0][0], CASES[2][0]]
    probs = []
    for kw, which in ((kws[0], "full"), (kws[1], "empty")):
        pr = P.anderson(**kw)
        b = Basis.enumerate(pr["NF"], pr["restrictions"])
        H = Op(pr["NF"], pr["H"])
        sell = Hop(H, b, mode="assembled")
        prod = Hop.product(H, b, group(pr, which, kw))
        v0 = torch.from_numpy(cvec(b.n, kw["seed"])).to(DEV)
        probs.append((b, sell, prod, v0))
    depths = [45, 38]
    mixed = Batch(8)
    ps = [mixed.add(prod, v0, d) for (b, sell, prod, v0), d in zip(probs, depths)]
    ss = [mixed.add(sell, v0, d) for (b, sell, prod, v0), d in zip(probs, depths)]
    for _ in range(7):
        mixed.run(8)
    for (b, sell, prod, v0), d, sp, sq in zip(probs, depths, ps, ss):
        st, a, bb = mixed.poll(sp)
        assert st == d
        ra, rb = sell.lanczos(v0, d)
        assert np.allclose(a, ra, rtol=1e-10, atol=1e-12)
        assert np.allclose(bb[:d], rb, rtol=1e-10, atol=0)
        alone = Batch(1)
        s1 = alone.add(prod, v0, d)
        alone.run(d)
        _, a1, b1 = alone.poll(s1)
        assert np.array_equal(a, a1) and np.array_equal(bb, b1)
        _, a2, b2 = mixed.poll(sq)
        only = Batch(1)
        s2 = only.add(sell, v0, d)
        only.run(d)
        _, a3, b3 = only.poll(s2)
        assert np.array_equal(a2, a3) and np.array_equal(b2, b3)
    # Ritz vectors: kept basis and second pass, canonical order
    b, sell, prod, v0 = probs[0]
    m = 60
    bt = Batch(4)
    sk = bt.add(prod, v0, m, keep_basis=True)
    sr = bt.add(sell, v0, m, keep_basis=True)
    bt.run(m)
    _, a, bb = bt.poll(sk)
    T = np.diag(a[1:m + 1]).astype(complex)
    for i in range(1, m):
        T[i, i - 1] = T[i - 1, i] = bb[i]
    w, Z = heev_host(T)
    z = (Z[:, 0] * np.exp(-1j * np.angle(Z[np.argmax(np.abs(Z[:, 0])), 0]))).real
    yp = bt.ritz(sk, z, b.zeros())
    ys = bt.ritz(sr, z, b.zeros())
    assert torch.linalg.vector_norm(yp - ys).item() < 1e-12 * torch.linalg.vector_norm(ys).item()
    s2 = bt.add(prod, v0, m, ycoef=z)
    bt.run(m)
    y2 = b.zeros()
    bt.poll(s2, want_coef=False, y=y2)
    assert torch.linalg.vector_norm(y2 - ys).item() < 1e-12 * torch.linalg.vector_norm(ys).item()
    # a Ritz vector's Rayleigh quotient is its Ritz value
    hy = sell.apply(yp.contiguous())
    rq = (torch.vdot(yp, hy) / torch.vdot(yp, yp)).item()
    assert abs(rq - w[0]) < 1e-9 * abs(w[0])
