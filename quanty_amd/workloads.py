"""Timed runs of the benchmark configurations on one GPU.

    python3 -m quanty_amd.workloads --config 0    # Ni2+ L2,3 XAS: 3 states x 3 dipoles, 200-step CF
    python3 -m quanty_amd.workloads --config 1    # 50-mode impurity: 3 lowest states, tol 1e-12
    python3 -m quanty_amd.workloads --config 2a   # 9 right-hand sides, 300-step block Lanczos
    python3 -m quanty_amd.workloads --config 2b   # 512 small impurity problems, batched vs one at a time
    python3 -m quanty_amd.workloads --config 3    # the serving stream at 10, 50, 200 and 500 req/s
    python3 -m quanty_amd.workloads --config 4    # 76-mode (two-word) L2,3 Green's function
    python3 -m quanty_amd.workloads --config layouts  # H*psi of the 50-mode impurity: SELL vs product layouts

Each run prints one JSON line: what was run, its sizes, wall times (HIP work
synchronized on both sides) and the derived rates. Accuracy against the oracle
is checked by the test suite, not here.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import default_context
from . import problems as P
from .engine import Basis, Hop, Op, block_cf_spectrum, cf_spectrum


def _log(msg):
    print(f"[workloads] {msg}", file=sys.stderr, flush=True)


def _sync_time(f, what: str = ""):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = f()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    if what:
        _log(f"{what}: {t:.3f} s")
    return out, t


def config0(reps: int = 5):
    """Ni2+ Oh charge-transfer L2,3 XAS through the CreateSpectra mirror: the 3
    lowest states of H_i (190 dets), the x/y/z dipoles into the 2p5 sector (120
    dets), 200-step Lanczos and the continued fraction on 2001 points."""
    from .api import CreateSpectra, Eigensystem, Operator

    ni = P.ni_xas()
    Hi, Hf = Operator(26, ni["Hi"]), Operator(26, ni["Hf"])
    Ts = [Operator(26, t) for t in ni["T"]]
    start = [26, 0, ["000000 1111111111 0000000000", 8, 8], ["111111 0000000000 1111111111", 16, 16]]
    opts = [["Emin", -10.0], ["Emax", 20.0], ["NE", 2000], ["Gamma", 0.4], ["NTri", 200], ["DenseBorder", 0]]

    def run():
        psis = Eigensystem(Hi, start, 3, [["Zero", 1e-14]])
        return CreateSpectra(Hf, Ts, psis, opts)

    def run_defaults():
        # the reference's defaults for spaces below DenseBorder: dense everywhere
        psis = Eigensystem(Hi, start, 3)
        return CreateSpectra(Hf, Ts, psis, opts[:-1])

    run()
    _, t = _sync_time(lambda: [run() for _ in range(reps)])
    t /= reps
    run_defaults()
    _, td = _sync_time(lambda: [run_defaults() for _ in range(reps)])
    td /= reps
    return {"config": "0", "what": "Ni2+ L2,3 XAS: Eigensystem (3 states, 190 dets) + CreateSpectra (3 dipoles x "
                                   "3 states, 120-det final sector, NTri 200, 2001 points, Gamma 0.4)",
            "seconds": t, "spectra": 9, "spectra_per_s": 9 / t,
            "seconds_dense_defaults": td, "spectra_per_s_dense_defaults": 9 / td}


def dense_small(seed: int = 5, reps: int = 5):
    """The dense small-problem path on a 1,410-determinant impurity (d + 1 bath
    shell, 16 electrons): ground state, one d<->bath transition, CreateSpectra by
    the dense method (closure, determinant matrix, Householder reduction,
    tridiagonal eigenproblem, poles) against 200-step Lanczos, 2001 points."""
    from .api import CreateSpectra, Operator, Wavefunction

    ctx = default_context()
    pr = P.anderson(seed=seed, nbath=1, nelec=16, full_baths=1, max_holes=1, max_elec=0, d_range=(0, 10))
    nf = pr["NF"]
    basis = Basis.enumerate(nf, pr["restrictions"], ctx)
    hop = Hop(Op(nf, pr["H"], ctx), basis)
    E, vecs, _, _ = hop.ground_states(1, tol=1e-14)
    psi = Wavefunction(nf, basis=basis, vec=vecs[0].clone())
    H = Operator(nf, dict(pr["H"]))
    bath = [m for b in pr["baths"] for m in b]
    T = Operator(nf, dict(P.random_one_body_transition(P.SplitMix64(seed), pr["dmodes"], bath)))
    opts = [["Emin", -10.0], ["Emax", 20.0], ["NE", 2000], ["Gamma", 0.4], ["NTri", 200],
            ["Restrictions", [nf, 0, *[[m, lo, hi] for m, lo, hi in pr["restrictions"]]]]]
    sd, _ = CreateSpectra(H, T, psi, opts)
    sl, _ = CreateSpectra(H, T, psi, opts + [["Method", "Lanczos"]])
    _, td = _sync_time(lambda: [CreateSpectra(H, T, psi, opts) for _ in range(reps)])
    _, tl = _sync_time(lambda: [CreateSpectra(H, T, psi, opts + [["Method", "Lanczos"]]) for _ in range(reps)])
    # stages of the dense method: matrix, then eigenvalues + projections (as
    # CreateSpectra uses), or every eigenvector (as KrylovBasis = true needs)
    from .engine import dense_poles
    from ._lib import QS_PARAM_HETRD, QS_PARAM_TRIDIAG_DC
    Hd, t_build = _sync_time(lambda: Hop(Op(nf, pr["H"], ctx), basis, mode="matfree").dense())

    def best_of(vectors, hetrd=1, times=3):
        # the reduction overwrites its matrix: a fresh copy per call; one
        # untimed call first (library kernels load on their first use)
        ctx.set_param(QS_PARAM_HETRD, hetrd)
        try:
            best = float("inf")
            for it in range(times + 1):
                A = Hd.clone()
                _, t = _sync_time(lambda: dense_poles(A, vecs[:1], ctx, vectors=vectors))
                if it:
                    best = min(best, t)
            return best
        finally:
            ctx.set_param(QS_PARAM_HETRD, 1)

    t_poles = best_of(False)
    # the same with rocSOLVER's zhetrd + Q^+ X instead of the cooperative reduction
    t_poles_rs = best_of(False, hetrd=0)
    t_eig = best_of(True)
    diff = float(np.max(np.abs(sd.G - sl.G)) / np.max(np.abs(sd.G)))
    # eigenvalues + projections of 3 vectors for random dense complex matrices:
    # the defaults (cooperative reduction up to 800 rows, tridiagonal divide and
    # conquer on the GPU from 256), rocSOLVER's zhetrd instead of the
    # cooperative reduction, and the host QL instead of the divide and conquer
    rng = np.random.default_rng(seed)
    reduction = []
    for m in (120, 190, 500, 800, 1000, 1410):
        B = rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))
        M = B + B.conj().T
        Xm = torch.from_numpy(rng.normal(size=(3, m)) + 1j * rng.normal(size=(3, m))).to(vecs.device)
        row = {"n": m}
        for hetrd, dc, key in ((1, 256, "ms_default"), (0, 256, "ms_rocsolver_zhetrd"), (1, 10 ** 9, "ms_host_ql")):
            ctx.set_param(QS_PARAM_HETRD, hetrd)
            ctx.set_param(QS_PARAM_TRIDIAG_DC, dc)
            try:
                best = None
                for _ in range(4):
                    Mt = torch.from_numpy(np.ascontiguousarray(M.T)).to(vecs.device).t()
                    _, t = _sync_time(lambda: dense_poles(Mt, Xm, ctx, vectors=False))
                    best = t if best is None else min(best, t)
            finally:
                ctx.set_param(QS_PARAM_HETRD, 1)
                ctx.set_param(QS_PARAM_TRIDIAG_DC, 256)
            row[key] = 1e3 * best
        reduction.append(row)
    # broadening of the spectrum: Gaussian 0.5, Lorentzian from an energy table
    table = [[-10.0, 0.2], [3.0, 0.2], [3.01, 0.6], [20.0, 0.6]]
    sd.Broaden(0.5, table)
    _, tb = _sync_time(lambda: [sd.Broaden(0.5, table) for _ in range(reps)])
    return {"config": "dense", "dim": basis.n, "seconds_dense": td / reps, "seconds_lanczos": tl / reps,
            "seconds_broaden_2001_points": tb / reps,
            "seconds_dense_matrix": t_build, "seconds_householder_poles": t_poles,
            "seconds_householder_poles_rocsolver_zhetrd": t_poles_rs,
            "seconds_all_eigenvectors": t_eig, "reduction_by_size": reduction,
            "lanczos200_vs_dense_max_rel_diff": diff}


def config1(seed: int = 1, nstates: int = 3, tol: float = 1e-12, max_krylov: int = 500):
    """The 3 lowest states of the 50-mode Anderson impurity (2,179,235 dets)."""
    ctx = default_context()
    pr = P.anderson(seed=seed)
    (b, hop), t_setup = _sync_time(lambda: (lambda b: (b, Hop(Op(pr["NF"], pr["H"], ctx), b, mode="assembled")))(
        Basis.enumerate(pr["NF"], pr["restrictions"], ctx)))
    (E, vecs, var, it), t = _sync_time(lambda: hop.ground_states(nstates, tol=tol, max_krylov=max_krylov),
                                       f"{nstates} lowest states")
    return {"config": "1", "what": f"{nstates} lowest states, tol {tol} on the variance, Krylov <= {max_krylov}",
            "dim": b.n, "setup_s": t_setup, "seconds": t, "E": [float(e) for e in E],
            "variance": [float(v) for v in var], "iterations": int(it)}, (pr, b, hop, vecs)


def config2a(seed: int = 1, nsteps: int = 300, state=None):
    """9 right-hand sides (3 random one-body d<->bath transitions x the 3 lowest
    states of the 50-mode impurity) in one 300-step block Lanczos, then the
    9x9 tensor continued fraction on 2001 points."""
    ctx = default_context()
    if state is None:
        _, state = config1(seed)
    pr, b, hop, vecs = state
    rng = P.SplitMix64(seed + 2000)
    bath = [m for bb in pr["baths"] for m in bb]
    Ts = [Hop(Op(pr["NF"], P.random_one_body_transition(rng, pr["dmodes"], bath), ctx), b, mode="matfree")
          for _ in range(3)]
    X = torch.stack([T.apply(vecs[j].contiguous()) for T in Ts for j in range(3)])
    (A, B, kept), t = _sync_time(lambda: hop.block_lanczos(X, nsteps))
    _, t_cf = _sync_time(lambda: block_cf_spectrum(A, B, kept, 0.0, -10.0, 20.0, 2000, 0.4, ctx))
    return {"config": "2a", "what": f"block Lanczos, block size 9 (3 T x 3 states), {nsteps} steps, then the "
                                    "tensor CF on 2001 points", "dim": b.n, "seconds_lanczos": t,
            "seconds_cf": t_cf, "kept_min": int(kept[1:].min()), "spectra": 9, "spectra_per_s": 9 / (t + t_cf)}


def config2b(n: int = 512, seed: int = 21, slots: int = 64):
    """512 impurity problems (d + 1..3 bath shells, dims log-uniform 10^2..10^5),
    each a ground state plus one 200-step spectrum: all submitted at once to the
    continuously batched server, and the same problems one at a time."""
    from . import serve

    ctx = default_context()
    warm = serve.request_stream(4, 0.0, seed=seed + 1, dim_lo=1e2, dim_hi=1e3, depth=(200, 200), nbath_max=3)
    serve.SpectrumServer(ctx, max_slots=4).serve(warm, realtime=False)
    reqs = serve.request_stream(n, 0.0, seed=seed, dim_lo=1e2, dim_hi=1e5, depth=(200, 200), nbath_max=3)
    t0 = time.perf_counter()
    served = serve.SpectrumServer(ctx, max_slots=slots, prep_workers=4).serve(reqs, realtime=False)
    t_served = time.perf_counter() - t0
    reqs2 = serve.request_stream(n, 0.0, seed=seed, dim_lo=1e2, dim_hi=1e5, depth=(200, 200), nbath_max=3)
    t0 = time.perf_counter()
    plain = serve.SpectrumServer(ctx, max_slots=1).serve(reqs2, realtime=False, plain=True)
    t_plain = time.perf_counter() - t0
    by = {r.rid: r for r in plain}
    same = all(np.array_equal(r.spectrum, by[r.rid].spectrum) for r in served)
    dims = np.array([r.dim for r in reqs])
    return {"config": "2b", "what": f"{n} problems, GS + 200-step spectrum each, batched ({slots} slots) vs one at "
                                    "a time", "dim_min": int(dims.min()), "dim_max": int(dims.max()),
            "dim_median": float(np.median(dims)), "seconds_batched": t_served, "seconds_plain": t_plain,
            "spectra_per_s_batched": n / t_served, "spectra_per_s_plain": n / t_plain,
            "speedup": t_plain / t_served, "identical_spectra": bool(same),
            "gs_steps_mean": float(np.mean([r.gs_steps for r in served]))}


def config3(requests: int = 64, rates=(10.0, 50.0, 200.0, 500.0), seed: int = 101, slots: int = 64):
    """The serving stream at several arrival rates: `requests` Poisson arrivals
    per rate (dims log-uniform 10^3..3*10^6, depth U{100..400}, Gamma
    U(0.1,0.5)), served with continuous batching and one at a time; latency
    from arrival to the spectrum on the host."""
    from . import serve

    ctx = default_context()
    warm = serve.request_stream(4, 0.0, seed=seed + 100, dim_lo=1e3, dim_hi=1e4)
    serve.SpectrumServer(ctx, max_slots=4).serve(warm, realtime=False)
    rows = []
    for rate in rates:
        reqs = serve.request_stream(requests, rate, seed=seed)
        served = serve.SpectrumServer(ctx, max_slots=slots).serve(reqs, realtime=True)
        _log(f"rate {rate}: served")
        plain = serve.SpectrumServer(ctx, max_slots=1).serve(serve.request_stream(requests, rate, seed=seed),
                                                              realtime=True, plain=True)
        _log(f"rate {rate}: plain")
        by = {r.rid: r for r in plain}
        s1, s2 = serve.latency_stats(served), serve.latency_stats(plain)
        rows.append({"rate_per_s": rate, "served": s1, "plain": s2,
                     "speedup": s1["spectra_per_s"] / s2["spectra_per_s"],
                     "identical_spectra": all(np.array_equal(r.spectrum, by[r.rid].spectrum) for r in served)})
    dims = np.array([r.dim for r in serve.request_stream(requests, 1.0, seed=seed)])
    return {"config": "3", "what": f"{requests} Poisson requests per rate, dims log-uniform 1e3..3e6, depth "
                                   f"U{{100..400}}, Gamma U(0.1,0.5), {slots} slots, 200 GB budget",
            "dim_min": int(dims.min()), "dim_max": int(dims.max()), "dim_median": float(np.median(dims)),
            "rates": rows}


def _spectrum_diff(a, b, a0, b0, ctx):
    g = cf_spectrum(a, b, 0.0, -10.0, 20.0, 2000, 0.4, ctx)
    g0 = cf_spectrum(a0, b0, 0.0, -10.0, 20.0, 2000, 0.4, ctx)
    return float(np.max(np.abs(g - g0)) / np.max(np.abs(g0)))


def layouts(seed: int = 1, reps: int = 30, nsteps: int = 200):
    """H*psi of the 50-mode Anderson impurity (2,179,235 determinants) in the
    assembled SELL layout and in the product layout for each restriction group
    as group A (and the automatic choice): one apply in the binding's own row
    order timed with HIP events over `reps` launches on the library's stream,
    and `nsteps` Lanczos steps from one start vector; the coefficients of every
    layout against SELL's."""
    ctx = default_context()
    dev = torch.device("cuda", ctx.device)
    pr = P.anderson(seed=seed)
    nf = pr["NF"]
    basis = Basis.enumerate(nf, pr["restrictions"], ctx)
    H = Op(nf, pr["H"], ctx)
    stream = torch.cuda.ExternalStream(ctx.lib.qs_ctx_stream(ctx.h), device=dev)
    g = np.random.default_rng(seed)
    x = torch.from_numpy(g.normal(size=basis.n) + 1j * g.normal(size=basis.n)).to(dev)

    def kernel_us(hop):
        order = torch.from_numpy(hop.row_order()).to(dev)
        xr = x[order].contiguous()
        yr = basis.zeros()
        hop.apply_rows(xr, yr)
        torch.cuda.synchronize()
        ctx.lib.qs_ctx_set_async(ctx.h, 1)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            ctx.check(ctx.lib.qs_hop_apply_rows(ctx.h, hop.h, xr.data_ptr(), yr.data_ptr()), "qs_hop_apply_rows")
        e1.record(stream)
        e1.synchronize()
        ctx.lib.qs_ctx_set_async(ctx.h, 0)
        return 1e3 * e0.elapsed_time(e1) / reps

    sell, t_sell = _sync_time(lambda: Hop(H, basis, mode="assembled"), "assemble SELL")
    (a0, b0), t_lz = _sync_time(lambda: sell.lanczos(x, nsteps), "SELL Lanczos")
    rows = [{"layout": "sell", "build_s": t_sell, "apply_us": kernel_us(sell), "lanczos_ms": 1e3 * t_lz,
             "stream_bytes": sell.stats()["stream_bytes"]}]
    del sell
    groups = [("auto", None), ("d", pr["dmodes"]), ("B1+B2", [m for b in pr["baths"][:2] for m in b]),
              ("B3+B4", [m for b in pr["baths"][2:] for m in b])]
    for name, am in groups:
        hop, t_b = _sync_time(lambda: Hop.product(H, basis, am), f"product layout, A = {name}")
        (a, b), t_lz = _sync_time(lambda: hop.lanczos(x, nsteps), f"product Lanczos, A = {name}")
        info = hop.product_info()
        rows.append({"layout": f"product A={name}", "group_a": hop.amodes, "build_s": t_b,
                     "apply_us": kernel_us(hop), "lanczos_ms": 1e3 * t_lz,
                     "stream_bytes": hop.stats()["stream_bytes"], "load_cost": info["load_cost"],
                     "groups": info["groups"], "slices": info["slices"], "entries": info["entries"],
                     "mapped": info["mapped"],
                     # the first 40 coefficients (later ones of a recurrence without
                     # re-orthogonalization follow rounding differences) and the
                     # broadened spectra of all nsteps
                     "coef40_max_rel_diff": float(max(np.max(np.abs(a[:41] - a0[:41])) / np.max(np.abs(a0[:41])),
                                                      np.max(np.abs(b[:40] - b0[:40])) / np.max(np.abs(b0[:40])))),
                     "spectrum_max_rel_diff": _spectrum_diff(a, b, a0, b0, ctx)})
        del hop
    return {"config": "layouts", "dim": basis.n, "lanczos_steps": nsteps, "rows": rows}


def product_apply(seed: int = 1, reps: int = 10):
    """The 50-mode impurity in the product layout with group A = B3+B4 (the
    automatic choice), applied `reps` times in its own row order: the kernel
    the product-layout counters are taken on (one kernel, one split)."""
    ctx = default_context()
    pr = P.anderson(seed=seed)
    nf = pr["NF"]
    basis = Basis.enumerate(nf, pr["restrictions"], ctx)
    hop = Hop.product(Op(nf, pr["H"], ctx), basis, [m for b in pr["baths"][2:] for m in b])
    g = np.random.default_rng(0)
    x = torch.from_numpy(g.normal(size=basis.n) + 1j * g.normal(size=basis.n)).to(basis.zeros().device)
    y = basis.zeros()
    _, t = _sync_time(lambda: [hop.apply_rows(x, y) for _ in range(reps)])
    return {"config": "product_apply", "dim": basis.n, "group_a": hop.amodes, "applies": reps,
            "us_per_apply": 1e6 * t / reps}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--config", required=True,
                    choices=["0", "1", "2a", "2b", "3", "4", "dense", "layouts", "product_apply"])
    ap.add_argument("--requests", type=int, default=64, help="requests per rate (config 3)")
    args = ap.parse_args()
    if args.config == "0":
        out = config0()
    elif args.config == "dense":
        out = dense_small()
    elif args.config == "1":
        out, _ = config1()
    elif args.config == "2a":
        out = config2a()
    elif args.config == "2b":
        out = config2b()
    elif args.config == "3":
        out = config3(args.requests)
    elif args.config == "layouts":
        out = layouts()
    elif args.config == "product_apply":
        out = product_apply()
    else:
        out = config4()
    print(json.dumps(out), flush=True)


def config4(seed: int = 1, nsteps: int = 300):
    """The 76-mode problem held as two-word determinants: ground state of H_i
    in the 2p6 sector, the L-edge dipole into the 2p5 sector, a 300-step Lanczos
    on H_f and the continued fraction."""
    ctx = default_context()
    pr = P.large76(seed=seed)
    nf = pr["NF"]
    (gs, t_gs_basis) = _sync_time(lambda: Basis.enumerate(nf, pr["gs_restrictions"], ctx), "enumerate 2p6")
    (fs, t_fs_basis) = _sync_time(lambda: Basis.enumerate(nf, pr["fs_restrictions"], ctx), "enumerate 2p5")
    (hi, t_hi) = _sync_time(lambda: Hop(Op(nf, pr["Hi"], ctx), gs, mode="assembled"), "assemble H_i")
    (E, vecs, var, it), t_gs = _sync_time(lambda: hi.ground_states(1, tol=1e-12, max_krylov=200), "ground state")
    hi_stats = hi.stats()
    del hi
    (hf, t_hf) = _sync_time(lambda: Hop(Op(nf, pr["Hf"], ctx), fs, mode="assembled"), "assemble H_f")
    T = Hop(Op(nf, pr["T"], ctx), fs, gs, mode="matfree")
    tpsi, t_t = _sync_time(lambda: T.apply(vecs[0].contiguous()), "T psi")
    (a, b), t_lz = _sync_time(lambda: hf.lanczos(tpsi, nsteps), f"{nsteps}-step Lanczos")
    _, t_cf = _sync_time(lambda: cf_spectrum(a, b, float(E[0]), -10.0, 20.0, 2000, 0.4, ctx), "continued fraction")
    st = hf.stats()
    return {"config": "4", "what": "76 modes (two-word determinants): GS of H_i, dipole into 2p5, "
                                   f"{nsteps}-step Lanczos on H_f, CF on 2001 points",
            "gs_dim": gs.n, "fs_dim": fs.n, "E0": float(E[0]), "gs_variance": float(var[0]), "gs_iterations": int(it),
            "seconds": {"enumerate_gs": t_gs_basis, "enumerate_fs": t_fs_basis, "assemble_hi": t_hi,
                        "ground_state": t_gs, "assemble_hf": t_hf, "T_psi": t_t, "lanczos": t_lz, "cf": t_cf},
            "hf_nnz_eff": st["nnz_eff"], "hf_stream_bytes": st["stream_bytes"], "hf_packed": st["packed"],
            "hi_nnz_eff": hi_stats["nnz_eff"], "lanczos_ms_per_step": 1e3 * t_lz / nsteps}


if __name__ == "__main__":
    main()
