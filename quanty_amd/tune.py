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

from ._lib import QS_PARAM_SHIFT_MIN, QS_PARAM_TERM_LDS, QSError
from .engine import Basis, Hop, Op, product_candidates

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


def time_apply_rows(hop: Hop, x: torch.Tensor, y: torch.Tensor, reps: int) -> float:
    """Average H*psi time in microseconds of a product-layout binding in its own
    row order (qs_hop_apply_rows, what the batch engine runs), over `reps`
    back-to-back launches."""
    ctx = hop.ctx
    stream = torch.cuda.ExternalStream(ctx.lib.qs_ctx_stream(ctx.h), device=x.device)
    torch.cuda.synchronize()
    ctx.lib.qs_ctx_set_async(ctx.h, 1)
    try:
        ctx.check(ctx.lib.qs_hop_apply_rows(ctx.h, hop.h, x.data_ptr(), y.data_ptr()), "qs_hop_apply_rows")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            ctx.check(ctx.lib.qs_hop_apply_rows(ctx.h, hop.h, x.data_ptr(), y.data_ptr()), "qs_hop_apply_rows")
        e1.record(stream)
        e1.synchronize()
    finally:
        ctx.lib.qs_ctx_set_async(ctx.h, 0)
    return e0.elapsed_time(e1) / reps * 1e3


def time_product(op: Op, basis: Basis, reps: int = 10, seed: int = 1):
    """(microseconds, group A modes, index in product_candidates) of the
    product layout with the split of fewest estimated loads, or (None, None,
    None) when no restriction gives one."""
    try:
        hop = Hop.product(op, basis)
    except QSError:
        return None, None, None
    cands = product_candidates(basis.nf, basis.restrictions or [])
    g = np.random.default_rng(seed)
    x = torch.from_numpy(g.normal(size=basis.n) + 1j * g.normal(size=basis.n)).cuda()
    y = basis.zeros()
    return time_apply_rows(hop, x, y, reps), hop.amodes, cands.index(hop.amodes)


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
    """The layout the server assembles each request with: for impurity shapes
    of the serving family near each of `dims`, the fastest shift-block
    threshold of `candidates` by timed H*psi, and the product layout (the
    split of fewest estimated loads, timed in its own row order) where it is
    faster than that; written to `out` as rows {dim, threshold, timings,
    product timing, layout}. layout_for() and product_for() pick the row
    nearest a request's dim on a log scale."""
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
        best_us = min(r["us"] for r in rs)
        pus, amodes, cand = time_product(op, basis, reps=reps, seed=seed)
        layout = "product" if pus is not None and pus < best_us else "sell"
        rows.append({"dim": int(basis.n), "shape": {k: list(v) if isinstance(v, tuple) else v for k, v in shape.items()},
                     "shift_min": int(best), "us": {str(r["shift_min"]): round(r["us"], 2) for r in rs},
                     "product_us": None if pus is None else round(pus, 2), "product_group_a": amodes,
                     "product_candidate": cand, "layout": layout})
        print(f"dim {basis.n}: threshold {best}  " + "  ".join(f"{r['shift_min']}: {r['us']:.1f} us" for r in rs)
              + (f"  product: {pus:.1f} us" if pus is not None else "  product: none") + f"  -> {layout}", flush=True)
    with open(out, "w") as f:
        json.dump({"what": "H*psi layout by basis size (shift-block threshold of the sliced ELL, or the product "
                           "layout where it is faster), fastest timed H*psi per impurity shape",
                   "candidates": list(candidates), "rows": rows}, f, indent=1)
    return rows


def _row_for(dim: int):
    global _TABLE
    if _TABLE is None:
        _TABLE = json.load(open(LAYOUT_TABLE))["rows"] if os.path.exists(LAYOUT_TABLE) else []
    if not _TABLE or dim < 1:
        return None
    return min(_TABLE, key=lambda r: abs(math.log(r["dim"]) - math.log(dim)))


def layout_for(dim: int) -> int | None:
    """The shift-block threshold for a basis of `dim` determinants from the
    fitted table (the measured shape nearest in log dim), or None without a
    table."""
    row = _row_for(dim)
    return None if row is None else int(row["shift_min"])


def product_for(dim: int) -> int | None:
    """For a basis of `dim` determinants at least as large as where the fitted
    table starts to have the product layout faster than the sliced ELL (the
    geometric mean of the first such row's dim and the row before it): the
    index, in engine.product_candidates order, of the restriction group that
    row used as group A; else None. Whether the product layout wins depends on
    the shape as well (the rows per group); the server checks that on the
    binding it builds."""
    global _TABLE
    _row_for(1)
    rows = sorted(_TABLE or [], key=lambda r: r["dim"])
    for k, r in enumerate(rows):
        if r.get("layout") == "product":
            start = math.sqrt(r["dim"] * rows[k - 1]["dim"]) if k > 0 else 0.0
            return r.get("product_candidate") if dim >= start else None
    return None


def main():
    from . import default_context
    from . import problems as P

    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--family", action="store_true",
                    help="fit the per-size layout table of the serving family (layout_table.json)")
    ap.add_argument("--out", default=LAYOUT_TABLE, help="where --family writes the table")
    args = ap.parse_args()
    if args.family:
        fit_family(reps=args.reps, seed=args.seed, out=args.out)
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
