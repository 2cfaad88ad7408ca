"""Benchmark: spectra/s of the plain (one-at-a-time) GPU Lanczos spectrum on
the 50-mode Anderson impurity (2,179,235 determinants), plus the H*psi kernel
against the HBM roofline and the oracle timed on the host cores.

A step is one spectrum: T|psi0> (random one-body d<->bath transition),
200 Lanczos steps on the device-resident assembled operator, and the
continued fraction on 2001 points (Gamma = 0.4). Inputs are resident in HBM
before the timed region starts.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--ntri", type=int, default=200)
    p.add_argument("--ne", type=int, default=2000)
    p.add_argument("--seed", type=int, default=1)
    p.add_argument("--hpsi-reps", type=int, default=20)
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="target length of the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--stream", type=int, default=128, help="requests in the served/plain stream sample")
    p.add_argument("--rate", type=float, default=200.0, help="Poisson arrival rate of the stream (req/s)")
    p.add_argument("--slots", type=int, default=64)
    p.add_argument("--no-serve", action="store_true")
    p.add_argument("--shift-min", type=int, default=None,
                   help="shift-block threshold of the assembled operator (library default when omitted)")
    return p.parse_args()


def bench_serving(args, ctx):
    """The same Poisson stream of impurity-spectrum requests served with
    continuous batching and with one request at a time."""
    from quanty_amd import serve

    reqs = serve.request_stream(args.stream, args.rate, seed=args.seed + 100)
    warm = serve.request_stream(4, 0.0, seed=args.seed + 200, dim_lo=1e3, dim_hi=1e4)
    serve.SpectrumServer(ctx, max_slots=4).serve(warm, realtime=False)
    t0 = time.perf_counter()
    served = serve.SpectrumServer(ctx, max_slots=args.slots).serve(reqs, realtime=True)
    t_served = time.perf_counter() - t0
    log(f"served stream done in {t_served:.1f} s")
    reqs2 = serve.request_stream(args.stream, args.rate, seed=args.seed + 100)
    t0 = time.perf_counter()
    plain = serve.SpectrumServer(ctx, max_slots=1).serve(reqs2, realtime=True, plain=True)
    t_plain = time.perf_counter() - t0
    log(f"plain stream done in {t_plain:.1f} s")
    same = all(np.array_equal(a.spectrum, b.spectrum) for a, b in
               zip(sorted(served, key=lambda r: r.rid), sorted(plain, key=lambda r: r.rid)))
    s1, s2 = serve.latency_stats(served), serve.latency_stats(plain)
    dims = np.array([r.dim for r in reqs])
    return {"served": s1, "plain": s2, "speedup": s1["spectra_per_s"] / s2["spectra_per_s"],
            "wall_s": {"served": t_served, "plain": t_plain}, "identical_spectra": bool(same),
            "stream": {"requests": args.stream, "rate_per_s": args.rate, "dim_min": int(dims.min()),
                       "dim_max": int(dims.max()), "dim_median": float(np.median(dims)), "depth": "U{100..400}",
                       "gamma": "U(0.1,0.5)", "ne_points": 2001, "gs": "two-pass Lanczos, 200 steps",
                       "slots": args.slots}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from quanty_amd import default_context
    from quanty_amd import problems as P
    from quanty_amd.engine import Basis, Hop, Op, cf_spectrum

    ctx = default_context()
    if args.shift_min is not None:
        from quanty_amd._lib import QS_PARAM_SHIFT_MIN
        ctx.set_param(QS_PARAM_SHIFT_MIN, args.shift_min)
    dev = torch.device("cuda", 0)
    pr = P.anderson(seed=args.seed)
    nf = pr["NF"]
    t0 = time.perf_counter()
    basis = Basis.enumerate(nf, pr["restrictions"], ctx)
    t_enum = time.perf_counter() - t0
    H = Op(nf, pr["H"], ctx)
    t0 = time.perf_counter()
    hop = Hop(H, basis, mode="assembled")
    t_asm = time.perf_counter() - t0
    st = hop.stats()
    rng = P.SplitMix64(args.seed + 1000)
    T = Op(nf, P.random_one_body_transition(rng, pr["dmodes"], [m for b in pr["baths"] for m in b]), ctx)
    thop = Hop(T, basis, mode="matfree")
    g = np.random.default_rng(args.seed)
    psi0 = torch.from_numpy(g.normal(size=basis.n) + 1j * g.normal(size=basis.n)).to(dev)
    psi0 /= torch.linalg.vector_norm(psi0)
    tpsi = basis.zeros()

    def spectrum():
        thop.apply(psi0, tpsi)
        a, b = hop.lanczos(tpsi, args.ntri)
        return cf_spectrum(a, b, 0.0, -10.0, 20.0, args.ne, 0.4, ctx)

    for _ in range(args.warmup):
        spectrum()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g_last = spectrum()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    assert np.all(np.isfinite(g_last))
    log(f"plain spectra on the 2.18M basis: {ms_per_step:.1f} ms each")

    # H*psi kernel time with HIP events on the library's stream
    stream = torch.cuda.ExternalStream(ctx.lib.qs_ctx_stream(ctx.h), device=dev)
    x = torch.from_numpy(g.normal(size=basis.n) + 1j * g.normal(size=basis.n)).to(dev)
    y = basis.zeros()
    torch.cuda.synchronize()
    ctx.lib.qs_ctx_set_async(ctx.h, 1)
    xp, yp = x.data_ptr(), y.data_ptr()
    ctx.check(ctx.lib.qs_hop_apply(ctx.h, hop.h, xp, yp), "qs_hop_apply")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.hpsi_reps):
        ctx.check(ctx.lib.qs_hop_apply(ctx.h, hop.h, xp, yp), "qs_hop_apply")
    e1.record(stream)
    e1.synchronize()
    ctx.lib.qs_ctx_set_async(ctx.h, 0)
    hpsi_ms = e0.elapsed_time(e1) / args.hpsi_reps
    n = basis.n

    # plain device copy of a known size: the practical HBM ceiling, and the
    # calibration of rocprof's FETCH_SIZE/WRITE_SIZE counters
    src = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)   # 1 GiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record()
    for _ in range(5):
        dst.copy_(src)
    c1.record()
    c1.synchronize()
    copy_ms = c0.elapsed_time(c1) / 5
    copy_gbps = 2 * src.numel() * 8 / (copy_ms * 1e-3) / 1e9
    del src, dst
    bytes_model = n * (8 + 16 + 16) + 16 * st["nnz_eff"]          # determinant + x_i + y_i + gathered x_j
    bytes_stream = st["stream_bytes"] + n * 16 * 2 + 16 * st["nnz_offdiag"]
    achieved = bytes_model / (hpsi_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pr, basis, args)
        log("cpu baseline done")
    srv = None if args.no_serve else bench_serving(args, ctx)

    out = {
        "metric": "spectra/s (plain one-at-a-time GPU Lanczos, 50-mode Anderson impurity, 2,179,235 dets)",
        "value": args.steps / elapsed,
        "unit": "spectra/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (fp64 complex)",
        "data": "synthetic (seeded Anderson impurity, SplitMix64 seed %d)" % args.seed,
        "config": {"workload": "anderson-50-modes-N28", "dim": n, "ntri": args.ntri, "ne_points": args.ne + 1,
                   "gamma": 0.4, "psi0": "random normalized vector", "transition": "random one-body d<->bath",
                   "parallelism": "single GPU"},
        "hpsi": {"ms": hpsi_ms, "nnz_eff": st["nnz_eff"], "nnz_stored": st["nnz_stored"],
                 "nnz_offdiag": st["nnz_offdiag"], "nblocks": st["nblocks"], "block_hits": st["block_hits"],
                 "ell_slots": st["ell_slots"], "packed": st["packed"], "shift_min": st["shift_min"],
                 "stream_bytes": st["stream_bytes"], "bytes_model": bytes_model, "bytes_stream_model": bytes_stream,
                 "GBps_model": achieved, "GBps_stream": bytes_stream / (hpsi_ms * 1e-3) / 1e9},
        "setup_s": {"enumerate": t_enum, "assemble": t_asm},
        "copy_1GiB": {"ms": copy_ms, "GBps": copy_gbps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_spmv<1,1,packed,lds> (assembled H*psi)"},
        "cpu_baseline": cpu,
        "serving": srv,
    }
    print(json.dumps(out))


def cpu_baseline(pr, basis, args):
    """The oracle's push-form H*psi (the reference algorithm: per (det, term)
    byte-wise ladder and hash insert) on a bounded sample of source rows, all
    host threads; spectra/s = 1 / (ntri * time of one full H*psi)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from quanty_amd import det as D

    nf = pr["NF"]
    threads = min(os.cpu_count() or 1, 16)
    op = oracle.Op(nf, pr["H"])
    words = basis.words()
    g = np.random.default_rng(0)

    def run(rows):
        sel = np.sort(g.choice(basis.n, rows, replace=False))
        psi = oracle.Vec.from_arrays(nf, D.words_to_bytes(words[sel], nf), np.ones(rows, dtype=complex))
        t0 = time.perf_counter()
        op.apply(psi, restrictions=pr["restrictions"], nthreads=threads)
        return time.perf_counter() - t0

    probe = 2000
    t = run(probe)
    rows = int(min(basis.n, max(probe, probe * args.cpu_seconds / max(t, 1e-3))))
    t = run(rows)
    t_full = t * basis.n / rows
    return {"value": 1.0 / (args.ntri * t_full), "unit": "spectra/s", "cores": threads, "kind": "port",
            "sample": f"oracle push-form H*psi over {rows} of {basis.n} source rows in {t:.2f} s "
                      f"({threads} threads); one H*psi extrapolated to {t_full:.2f} s, x{args.ntri} Lanczos steps"}


if __name__ == "__main__":
    main()
