"""Autotuner for the assembled operator layout.

The assembled H*psi stores, per 64-row slice, shift blocks (one 16-byte header
for a term whose rows read consecutive columns) and per-entry ELL slots. Which
groups become blocks is set by QS_PARAM_SHIFT_MIN; the best value depends on
the operator's structure. `tune_shift_min` assembles the operator for each
candidate, times H*psi with HIP events on the library's stream and returns the
fastest (optionally leaving it set on the context). `tune_term_lds` does the
same for the largest term table (values and parity masks) staged in LDS:
staging saves global reads per entry but takes LDS from other workgroups.

    python3 -m quanty_amd.tune            # the 50-mode Anderson impurity, 2,179,235 dets
    python3 -m quanty_amd.tune --family   # the table the server assembles requests with
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np
import torch

from ._lib import QS_PARAM_SHIFT_MIN, QS_PARAM_TERM_LDS
from .engine import Basis, Hop, Op

CANDIDATES = (65, 64, 56, 48, 32, 24, 16, 8)
LDS_CANDIDATES = (0, 24 * 1024, 64 * 1024)


def time_apply(hop: Hop, x: torch.Tensor, y: torch.Tensor, reps: int) -> float:
    """Average H*psi time in microseconds over `reps` back-to-back launches."""
    ctx = hop.ctx
    stream = torch.cuda.ExternalStream(ctx.lib.qs_ctx_stream(ctx.h), device=x.device)
    torch.cuda.synchronize()
    ctx.lib.qs_ctx_set_async(ctx.h, 1)
    try:
        ctx.check(ctx.lib.qs_hop_apply(ctx.h, hop.h, x.data_ptr(), y.data_ptr()), "qs_hop_apply")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            ctx.check(ctx.lib.qs_hop_apply(ctx.h, hop.h, x.data_ptr(), y.data_ptr()), "qs_hop_apply")
        e1.record(stream)
        e1.synchronize()
    finally:
        ctx.lib.qs_ctx_set_async(ctx.h, 0)
    return e0.elapsed_time(e1) / reps * 1e3


def tune_shift_min(op: Op, basis: Basis, candidates=CANDIDATES, reps: int = 10, apply: bool = True,
                   seed: int = 1):
    """Time H*psi of `op` on `basis` for each shift-block threshold.

    Returns (best threshold, rows) with one dict per candidate: threshold,
    microseconds, blocks, entries in blocks, ELL slots, stream bytes. With
    apply=True the best threshold stays set on the operator's context."""
    ctx = op.ctx
    old = ctx.get_param(QS_PARAM_SHIFT_MIN)
    g = np.random.default_rng(seed)
    x = torch.from_numpy(g.normal(size=basis.n) + 1j * g.normal(size=basis.n)).cuda()
    y = basis.zeros()
    rows = []
    try:
        for sm in candidates:
            ctx.set_param(QS_PARAM_SHIFT_MIN, sm)
            hop = Hop(op, basis, mode="assembled")
            st = hop.stats()
            us = time_apply(hop, x, y, reps)
            rows.append({"shift_min": sm, "us": us, "nblocks": st["nblocks"], "block_hits": st["block_hits"],
                         "ell_slots": st["ell_slots"], "stream_bytes": st["stream_bytes"]})
            del hop
    finally:
        ctx.set_param(QS_PARAM_SHIFT_MIN, old)
    best = min(rows, key=lambda r: r["us"])["shift_min"]
    if apply:
        ctx.set_param(QS_PARAM_SHIFT_MIN, best)
    return best, rows


def tune_term_lds(op: Op, basis: Basis, candidates=LDS_CANDIDATES, reps: int = 10, apply: bool = True,
                  seed: int = 1):
    """Time H*psi of `op` on `basis` for each limit on the term tables staged
    in LDS (0: always read from global memory). The operator is assembled once;
    the limit takes effect at launch. Returns (best limit, rows); with
    apply=True the best limit stays set on the operator's context."""
    ctx = op.ctx
    old = ctx.get_param(QS_PARAM_TERM_LDS)
    g = np.random.default_rng(seed)
    x = torch.from_numpy(g.normal(size=basis.n) + 1j * g.normal(size=basis.n)).cuda()
    y = basis.zeros()
    hop = Hop(op, basis, mode="assembled")
    rows = []
    try:
        for lds in candidates:
            ctx.set_param(QS_PARAM_TERM_LDS, lds)
            rows.append({"term_lds": lds, "us": time_apply(hop, x, y, reps)})
    finally:
        ctx.set_param(QS_PARAM_TERM_LDS, old)
    best = min(rows, key=lambda r: r["us"])["term_lds"]
    if apply:
        ctx.set_param(QS_PARAM_TERM_LDS, best)
    return best, rows


LAYOUT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "layout_table.json")
_TABLE = None


def fit_family(dims=(1e3, 1e4, 1e5, 5e5, 1.5e6, 3e6), candidates=(65, 56, 48, 32), reps: int = 10,
               seed: int = 1, out: str = LAYOUT_TABLE):
    """The shift-block threshold the server assembles each request with: for
    impurity shapes of the serving family near each of `dims`, the fastest
    threshold of `candidates` by timed H*psi; written to `out` as rows
    {dim, threshold, timings}. layout_for() picks the row nearest a
    request's dim on a log scale."""
    from . import default_context
    from . import problems as P
    from . import serve

    ctx = default_context()
    rows = []
    for target in dims:
        dim, shape = serve.pick_shape(target)
        pr = P.anderson(seed=seed, **shape)
        basis = Basis.enumerate(pr["NF"], pr["restrictions"], ctx)
        op = Op(pr["NF"], pr["H"], ctx)
        best, rs = tune_shift_min(op, basis, candidates=candidates, reps=reps, apply=False, seed=seed)
        rows.append({"dim": int(basis.n), "shape": {k: list(v) if isinstance(v, tuple) else v for k, v in shape.items()},
                     "shift_min": int(best), "us": {str(r["shift_min"]): round(r["us"], 2) for r in rs}})
        print(f"dim {basis.n}: threshold {best}  " + "  ".join(f"{r['shift_min']}: {r['us']:.1f} us" for r in rs),
              flush=True)
    with open(out, "w") as f:
        json.dump({"what": "shift-block threshold by basis size, fastest timed H*psi per impurity shape",
                   "candidates": list(candidates), "rows": rows}, f, indent=1)
    return rows


def layout_for(dim: int) -> int | None:
    """The shift-block threshold for a basis of `dim` determinants from the
    fitted table (the measured shape nearest in log dim), or None without a
    table."""
    global _TABLE
    if _TABLE is None:
        _TABLE = json.load(open(LAYOUT_TABLE))["rows"] if os.path.exists(LAYOUT_TABLE) else []
    if not _TABLE or dim < 1:
        return None
    row = min(_TABLE, key=lambda r: abs(math.log(r["dim"]) - math.log(dim)))
    return int(row["shift_min"])


def main():
    from . import default_context
    from . import problems as P

    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--family", action="store_true",
                    help="fit the per-size threshold table of the serving family (layout_table.json)")
    args = ap.parse_args()
    if args.family:
        fit_family(reps=args.reps, seed=args.seed)
        return
    ctx = default_context()
    pr = P.anderson(seed=args.seed)
    t0 = time.perf_counter()
    basis = Basis.enumerate(pr["NF"], pr["restrictions"], ctx)
    op = Op(pr["NF"], pr["H"], ctx)
    best, rows = tune_shift_min(op, basis, reps=args.reps, apply=False, seed=args.seed)
    print(f"50-mode Anderson impurity, {basis.n} determinants, {len(pr['H'])} terms "
          f"({time.perf_counter() - t0:.1f} s); fastest threshold {best}")
    print("| threshold | H*psi us | blocks | entries in blocks | ELL slots | stream MB |")
    print("|---|---|---|---|---|---|")
    for r in rows:
        print(f"| {r['shift_min']} | {r['us']:.1f} | {r['nblocks']:,} | {r['block_hits']:,} | {r['ell_slots']:,} | "
              f"{r['stream_bytes'] / 1e6:.1f} |")
    lbest, lrows = tune_term_lds(op, basis, reps=args.reps, apply=False, seed=args.seed)
    print(f"\nlimit on the term tables staged in LDS, layout threshold {ctx.get_param(QS_PARAM_SHIFT_MIN)} "
          f"({op.nterms} terms); fastest {lbest}")
    print("| LDS limit (bytes) | H*psi us |")
    print("|---|---|")
    for r in lrows:
        print(f"| {r['term_lds']} | {r['us']:.1f} |")


if __name__ == "__main__":
    main()
