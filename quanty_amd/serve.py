"""Continuous-batching spectrum service.

A request is one impurity spectrum: a seeded Anderson impurity (d shell plus
bath shells, sizes from 10^3 to a few 10^6 determinants), its ground state
(Lanczos from a seeded random vector weighted towards low-energy determinants,
extended in steps until the lowest Ritz pair has
|H psi - E psi|^2 <= tol max(1, E^2); the Ritz vector is formed from the
Lanczos vectors kept on the device, or, with keep_basis=False, rebuilt by a
second pass from the same start), T|psi0> for a random one-body d<->bath
transition T, a Lanczos tridiagonalization of the requested depth, and the
continued fraction on the grid. Requests are admitted as slots and HBM budget free up; every admitted
request is one slot of a `Batch`, so all of them advance together, one set of
launches per Lanczos step. Bases, operators and vectors stay on the device
from admission to completion. The plain baseline runs the same requests with
one slot, one at a time.
"""
from __future__ import annotations

import json
import math
import os
import queue
import sys
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import problems as P
from ._lib import QS_PARAM_ALLOC_TOP, QS_PARAM_SHIFT_MIN, QSError
from .engine import Basis, Batch, Hop, Op, cf_spectrum, product_candidates, tridiag_lowest
from .tune import layout_for, product_for

# (nbath, full_baths, nelec, d_range, max_holes, max_elec) -> sector size, cached
_FAMILY = None


_FAMILY_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "impurity_family.json")


def impurity_family():
    """Anderson-impurity shapes with their sector sizes, from about 10^2 to 10^8.
    The table ships as impurity_family.json next to this module (written by
    `python -m quanty_amd.serve --write-family`); without it the sizes are
    counted here, in memory."""
    global _FAMILY
    if _FAMILY is not None:
        return _FAMILY
    if os.path.exists(_FAMILY_FILE):
        data = json.load(open(_FAMILY_FILE))
        _FAMILY = [(int(n), dict(kw, d_range=tuple(kw["d_range"]))) for n, kw in data]
        return _FAMILY
    _FAMILY = _count_family()
    return _FAMILY


def _count_family():
    out = []
    for nbath in (1, 2, 3, 4):
        for full in range(0, nbath + 1):
            for nd in (6, 7, 8, 9):
                nelec = 10 * full + nd
                for mh in range(0, 4):
                    for me in range(0, 4):
                        if full == 0 and mh:
                            continue
                        if full == nbath and me:
                            continue
                        for dw in (1, 2):
                            kw = dict(nbath=nbath, full_baths=full, nelec=nelec, d_range=(max(0, nd - dw), min(10, nd + dw)),
                                      max_holes=mh, max_elec=me)
                            n = Basis.count(10 * (1 + nbath), P.anderson_restrictions(**kw))
                            if n >= 50:
                                out.append((n, kw))
    out.sort(key=lambda t: t[0])
    return out


_SHAPES: dict = {}


def pick_shape(target_dim: float, nbath_max: int = 4):
    """The impurity shape (at most nbath_max bath shells) whose sector size is
    closest to target_dim on a log scale (the first such in family order on a tie)."""
    if nbath_max not in _SHAPES:
        fam = [t for t in impurity_family() if t[1]["nbath"] <= nbath_max]
        _SHAPES[nbath_max] = (fam, np.log(np.array([t[0] for t in fam], dtype=float)))
    fam, logs = _SHAPES[nbath_max]
    return fam[int(np.argmin(np.abs(logs - math.log(target_dim))))]


@dataclass(order=True)
class Request:
    arrival: float
    rid: int = field(compare=False)
    seed: int = field(compare=False)
    shape: dict = field(compare=False, default_factory=dict)
    dim: int = field(compare=False, default=0)
    depth: int = field(compare=False, default=200)
    gamma: float = field(compare=False, default=0.3)
    # results
    t_admit: float = field(compare=False, default=-1.0)
    t_done: float = field(compare=False, default=-1.0)
    E0: float = field(compare=False, default=0.0)
    gs_residual: float = field(compare=False, default=0.0)
    gs_steps: int = field(compare=False, default=0)
    spectrum: np.ndarray | None = field(compare=False, default=None)
    gs_state: np.ndarray | None = field(compare=False, default=None, repr=False)
    rejected: str | None = field(compare=False, default=None)  # why the server refused it
    a: np.ndarray | None = field(compare=False, default=None)
    b: np.ndarray | None = field(compare=False, default=None)
    layout: str = field(compare=False, default="")  # "sell" or "product": how its H was bound
    # the request's payload: NF, H and T terms, restrictions (made from seed and
    # shape when absent)
    problem: dict | None = field(compare=False, default=None, repr=False)


def make_problem(seed: int, shape: dict) -> dict:
    """The seeded impurity of a request: H, its sector restrictions and a random
    one-body d<->bath transition T."""
    pr = P.anderson(seed=seed, **shape)
    rng = P.SplitMix64(seed ^ 0x5DEECE66D)
    bath = [m for b in pr["baths"] for m in b]
    return {"NF": pr["NF"], "H": pr["H"], "restrictions": pr["restrictions"],
            "T": P.random_one_body_transition(rng, pr["dmodes"], bath)}


def request_stream(n: int, rate: float, seed: int = 7, dim_lo=1e3, dim_hi=3e6, depth=(100, 400),
                   gamma=(0.1, 0.5), nbath_max: int = 4, payload: bool = True) -> list[Request]:
    """Poisson arrivals at `rate` req/s (all at t = 0 when rate is 0); dim
    log-uniform; depth U{lo..hi}; Gamma U(lo, hi); at most nbath_max bath shells.
    payload: each request carries its operator terms, as a client's request
    would (made here, before any clock starts)."""
    rng = P.SplitMix64(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += -math.log(max(rng.uniform(), 1e-300)) / rate if rate > 0 else 0.0
        target = math.exp(rng.uniform(math.log(dim_lo), math.log(dim_hi)))
        dim, shape = pick_shape(target, nbath_max)
        d = depth[0] + int(rng.uniform() * (depth[1] - depth[0] + 1))
        g = rng.uniform(*gamma)
        out.append(Request(arrival=t, rid=i, seed=int(rng.next() & 0x7FFFFFFF), shape=shape, dim=dim, depth=d, gamma=g))
    if payload:
        for r in out:
            r.problem = make_problem(r.seed, r.shape)
    return out


class PatternCache:
    """Assembled connection streams by operator structure, shared by requests
    whose Hamiltonians differ only in their values (the same impurity shape
    with another seed): such a request binds its operator to a cached stream
    with qs_hop_create_like instead of enumerating its basis and assembling.
    Keyed by (modes, restrictions, the ladder codes of H's terms in order), so
    a hit is exact; the library checks the compiled terms again and refuses a
    mismatch (a coefficient that vanishes drops a term), in which case the
    request is assembled on its own. Least recently used entries go once the
    streams exceed `cap_bytes`; the server counts the cached bytes against its
    HBM budget. A stream stays on the device while any request uses it.
    Process-wide: preparation workers and successive servers share it."""

    def __init__(self, cap_bytes: float = 24e9):
        self.cap = cap_bytes
        self.lock = threading.Lock()
        self.entries: "dict[tuple, tuple]" = {}  # key -> (basis, template hop, bytes), in use order
        self.bytes = 0
        self.hits = 0
        self.misses = 0

    @staticmethod
    def key(pr) -> tuple:
        return (pr["NF"], tuple((str(m), int(a), int(b)) for m, a, b in pr["restrictions"]),
                tuple(tuple(c) for c, _ in pr["H"]))

    def get(self, key):
        with self.lock:
            e = self.entries.pop(key, None)
            if e is None:
                self.misses += 1
                return None
            self.entries[key] = e  # most recently used last
            self.hits += 1
            return e[0], e[1]

    def put(self, key, basis, hop, nbytes: int):
        if nbytes > self.cap:
            return
        with self.lock:
            if key in self.entries:
                return
            self.entries[key] = (basis, hop, nbytes)
            self.bytes += nbytes
            while self.bytes > self.cap and self.entries:
                k = next(iter(self.entries))
                self.bytes -= self.entries.pop(k)[2]

    def clear(self):
        with self.lock:
            self.entries.clear()
            self.bytes = 0
            self.hits = self.misses = 0


PATTERNS = PatternCache()

# the product layout is kept only when its groups average at least this many rows
PRODUCT_MIN_GROUP_ROWS = 128


class _Job:
    """Device-resident state of an admitted request."""

    def __init__(self, req: Request, ctx, gs_depth: int, start_temperature: float | None = 0.3,
                 start_floor: float = 1e-3, patterns: PatternCache | None = PATTERNS, product: bool = False):
        self.req = req
        pr = req.problem if req.problem is not None else make_problem(req.seed, req.shape)
        self.nf = pr["NF"]
        self.H = Op(self.nf, pr["H"], ctx)
        key = PatternCache.key(pr) if patterns is not None else None
        hit = patterns.get(key) if patterns is not None else None
        self.hop = None
        self.shared = False
        if hit is not None:
            try:
                self.hop = Hop.like(self.H, hit[1])
                self.basis = hit[0]
                self.shared = True
            except QSError:
                self.hop = None
        if self.hop is None:
            # a basis and stream the cache will keep go to the top of the
            # library's device memory, away from the requests' short-lived blocks
            if patterns is not None:
                ctx.set_param(QS_PARAM_ALLOC_TOP, 1)
            try:
                self.basis = Basis.enumerate(self.nf, pr["restrictions"], ctx)
                # the layout the autotuner fitted for this basis size
                # (tune.fit_family): the product layout where it was faster,
                # else the sliced ELL with the fitted shift-block threshold
                cand = product_for(self.basis.n) if product else None
                if cand is not None:
                    groups = product_candidates(self.nf, pr["restrictions"])
                    try:
                        h = Hop.product(self.H, self.basis, groups[cand] if cand < len(groups) else None)
                        # groups of fewer rows than two wavefronts leave lanes idle that cost
                        # more than the layout saves (a 56-row group set measured slower)
                        if self.basis.n >= PRODUCT_MIN_GROUP_ROWS * h.product_info()["groups"]:
                            self.hop = h
                    except QSError:
                        self.hop = None
                if self.hop is None:
                    sm = layout_for(self.basis.n)
                    if sm is not None:
                        ctx.set_param(QS_PARAM_SHIFT_MIN, sm)
                    self.hop = Hop(self.H, self.basis, mode="assembled")
            finally:
                if patterns is not None:
                    ctx.set_param(QS_PARAM_ALLOC_TOP, 0)
            if patterns is not None:
                patterns.put(key, self.basis, self.hop, self.hop.stats()["stream_bytes"])
        self.T = Op(self.nf, pr["T"], ctx)
        req.layout = "product" if getattr(self.hop, "amodes", None) else "sell"
        self.thop = Hop(self.T, self.basis, mode="matfree")
        g = torch.Generator(device="cuda").manual_seed(req.seed)
        dev = torch.device("cuda", ctx.device)
        v = torch.randn(self.basis.n, dtype=torch.float64, device=dev, generator=g) + \
            1j * torch.randn(self.basis.n, dtype=torch.float64, device=dev, generator=g)
        if start_temperature is not None:
            # weight determinants by exp(-(H_ii - min H_ii)/T), at least start_floor:
            # the ground state lives mostly on low-energy determinants, and the
            # floor keeps every determinant (every symmetry sector) in the start
            d = self.hop.diagonal().real
            v = v * torch.clamp(torch.exp(-(d - d.min()) / start_temperature), min=start_floor)
        self.v0 = v.to(torch.complex128).contiguous()
        self.gs_depth = min(gs_depth, self.basis.n)
        self.phase = "gs1"
        self.gs_prev = None  # (lowest Ritz value, residual) of the last ground-state check
        self.restarts = 0  # ground-state restarts from the Ritz vector when the kept basis is full
        self.steps_done = 0  # ground-state steps of the runs before the last restart
        self.slot = -1
        # device bytes the request holds: its vectors, and the assembled stream
        # unless it shares a cached one (then only its own diagonal)
        own = 8 * self.basis.n if self.shared else self.hop.stats()["stream_bytes"]
        self.bytes = own + 6 * 16 * self.basis.n
        self.kept_bytes = 0  # Lanczos vectors kept during the ground-state phase

    def release(self):
        for k in ("hop", "thop", "H", "T", "basis", "v0", "gs_vector"):
            setattr(self, k, None)


# Library contexts of the preparation workers, kept for the life of the
# process. Each context owns a HIP stream, and a process has few hardware
# queues (4 by default): streams created anew for every server would come to
# share the serving stream's queue and serialize behind its work.
_WORKER_CTXS: dict[int, list] = {}
_WORKER_LOCK = threading.Lock()


def _worker_contexts(device: int, n: int) -> list:
    from ._lib import Context
    with _WORKER_LOCK:
        pool = _WORKER_CTXS.setdefault(device, [])
        while len(pool) < n:
            pool.append(Context(device))
        return pool[:n]


class _Prep:
    """Builds the state of admitted requests (basis, operators, assembled
    stream, start vector) on worker threads, each with its own library context,
    so that the preparation overlaps the batch's GPU steps."""

    def __init__(self, device: int, gs_depth: int, workers: int = 1, start=(0.3, 1e-3), patterns=None,
                 product: bool = False):
        self.gs_depth = gs_depth
        self.start = start
        self.patterns = patterns
        self.product = product
        self.todo: queue.Queue = queue.Queue()
        self.ready: queue.Queue = queue.Queue()
        self.ctxs = _worker_contexts(device, max(1, workers))
        self.threads = [threading.Thread(target=self._run, args=(c,), daemon=True) for c in self.ctxs]
        for t in self.threads:
            t.start()

    def _run(self, ctx):
        while True:
            r = self.todo.get()
            if r is None:
                return
            try:
                job = _Job(r, ctx, self.gs_depth, *self.start, patterns=self.patterns, product=self.product)
                torch.cuda.current_stream().synchronize()  # the start vector; library calls are synchronous
                self.ready.put(job)
            except BaseException as e:  # handed to the serving loop, which raises it
                self.ready.put(e)

    def close(self):
        for _ in self.threads:
            self.todo.put(None)
        for t in self.threads:
            t.join()


def estimate_bytes(dim: int, kept_vectors: int = 0) -> int:
    """Device bytes of a request: vectors, assembled stream, kept Lanczos vectors."""
    return dim * (6 * 16 + 8 * 110 + 16 * kept_vectors)


class SpectrumServer:
    """Admits requests into a `Batch` by slot and HBM budget and advances all
    admitted requests together."""

    def __init__(self, ctx, max_slots: int = 64, hbm_budget: float = 200e9, gs_depth: int = 60,
                 gs_max: int = 500, gs_step: int = 16, gs_tol: float = 1e-12, gs_restarts: int = 8,
                 chunk: int = 8, ne: int = 2000,
                 emin: float = -10.0, emax: float = 20.0, keep_basis: bool = True, prepare_ahead: bool = True,
                 prep_workers: int = 2, start_temperature: float | None = 0.3, start_floor: float = 1e-3,
                 share_patterns: bool = True, product_layout: bool = False):
        self.ctx = ctx
        # product_layout: requests the fitted layout table (tune.product_for) and
        # their group sizes favour are bound in the product layout. Off by
        # default: on the 10,000-request stream it served 52.8 spectra/s against
        # 53.4 with the sliced ELL alone, although a lone 2.18 M-row slot steps
        # 8.6% faster in it
        self.product_layout = product_layout
        # requests of one operator structure share one assembled stream (PatternCache)
        self.patterns = PATTERNS if share_patterns else None
        self.max_slots = max_slots
        self.budget = hbm_budget
        self.gs_depth = gs_depth
        self.gs_max = gs_max
        self.gs_restarts = gs_restarts
        self.gs_step = gs_step
        self.gs_tol = gs_tol
        self.chunk = chunk
        self.ne, self.emin, self.emax = ne, emin, emax
        self.keep_basis = keep_basis
        self.prepare_ahead = prepare_ahead
        self.prep_workers = prep_workers
        # ground-state start: random, weighted by exp(-(H_ii - min)/T) (at least
        # start_floor) unless start_temperature is None
        self.start_temperature = start_temperature
        self.start_floor = start_floor
        self.batch = Batch(max_slots, ctx)
        self.jobs: dict[int, _Job] = {}
        self.used = 0
        # requests whose normalized ground state is copied to the host
        # (Request.gs_state) for an independent check of the result
        self.keep_states: set = set()
        self.rejected: list[Request] = []  # requests refused by the budget

    def reserve(self, extra: float = 16e9):
        """Take the HBM budget (plus `extra` for the preparation threads'
        temporaries) from the driver once, as one chunk of the library's
        device memory, so that requests are placed in memory the process
        already holds: the driver clears memory it hands out, tens of
        milliseconds per GB, and serializes other threads' allocations
        meanwhile."""
        import ctypes as C
        p = C.c_void_p()
        self.ctx.check(self.ctx.lib.qs_dev_alloc(self.ctx.h, int(self.budget + extra), C.byref(p)), "qs_dev_alloc")
        self.ctx.check(self.ctx.lib.qs_dev_free(self.ctx.h, p), "qs_dev_free")

    def _advance(self, job: _Job, coeffs=None):
        """Move a job whose slot reached its depth to its next phase. coeffs:
        (completed steps, a, b) of its slot when already fetched. Returns True
        when the request is done."""
        r = job.req
        if job.phase == "gs1":
            m, a, b = coeffs if coeffs is not None else self.batch.poll(job.slot)
            # the previous check's Ritz value bounds the new lowest one from
            # above, and lies within its residual of an eigenvalue of H
            lo = hi = None
            if job.gs_prev is not None:
                E0, r0 = job.gs_prev
                hi = E0 + 1e-12 * max(1.0, abs(E0))
                lo = E0 - 1.01 * r0 - 1e-12 * max(1.0, abs(E0))
            E, z = tridiag_lowest(a[1:m + 1], b[1:m], lo, hi)
            res = float(abs(b[m] * z[m - 1]))
            job.gs_prev = (E, res)
            converged = res ** 2 <= self.gs_tol * max(1.0, E ** 2)
            if not converged and m < job.gs_max:
                job.gs_depth = min(job.gs_max, m + self.gs_step)
                job.extend = True  # the caller sets the new depth (one launch for all such jobs)
                return False
            if not converged and self.keep_basis and job.restarts < self.gs_restarts:
                # the kept basis is full: restart the Lanczos run from the Ritz vector
                y = self.batch.ritz(job.slot, z, job.basis.zeros())
                self.batch.remove(job.slot)
                y /= torch.linalg.vector_norm(y)
                job.restarts += 1
                job.steps_done += m
                job.gs_prev = None
                job.gs_depth = min(self.gs_depth, job.gs_max)
                job.slot = self.batch.add(job.hop, y.contiguous(), job.gs_depth, max_depth=job.gs_max,
                                          keep_basis=True)
                return False
            r.E0 = float(E)
            r.gs_residual = res
            r.gs_steps = job.steps_done + m
            if not self.keep_basis:
                self.batch.remove(job.slot)
                job.slot = self.batch.add(job.hop, job.v0, m, ycoef=z)
                job.gs_depth = m
                job.phase = "gs2"
                return False
            y = self.batch.ritz(job.slot, z, job.basis.zeros())
            self.batch.remove(job.slot)
            self.used -= job.kept_bytes
            job.kept_bytes = 0
        if job.phase in ("gs1", "gs2"):
            if job.phase == "gs2":
                y = job.basis.zeros()
                self.batch.poll(job.slot, want_coef=False, y=y)
                self.batch.remove(job.slot)
            y /= torch.linalg.vector_norm(y)
            job.gs_vector = y  # the normalized ground state (device)
            if r.rid in self.keep_states:
                r.gs_state = y.cpu().numpy()
            tpsi = job.thop.apply(y.contiguous(), ctx=self.ctx)
            job.slot = self.batch.add(job.hop, tpsi, r.depth)
            job.phase = "spec"
            return False
        _, a, b = coeffs if coeffs is not None else self.batch.poll(job.slot)
        self.batch.remove(job.slot)
        r.a, r.b = a[:r.depth + 1].copy(), b[:r.depth].copy()
        r.spectrum = cf_spectrum(r.a, r.b, r.E0, self.emin, self.emax, self.ne, r.gamma, self.ctx)[0]
        return True

    def _step_jobs(self, t0: float, done: list):
        """After a run: advance every job whose slot reached its depth. The
        coefficients of all of them come in one transfer and the depths of the
        ground states that need more steps are set with one launch."""
        steps = self.batch.steps(self.max_slots)
        ready = []
        for job in self.jobs.values():
            depth = job.gs_depth if job.phase in ("gs1", "gs2") else job.req.depth
            if steps[job.slot] >= depth:
                ready.append(job)
        need = [j for j in ready if j.phase in ("gs1", "spec")]
        fetched = {}
        if need:
            ld = max((j.gs_depth if j.phase == "gs1" else j.req.depth) for j in need) + 1
            st, A, Bc = self.batch.coeffs([j.slot for j in need], ld)
            for j, m, a, b in zip(need, st, A, Bc):
                fetched[j.req.rid] = (int(m), a, b)
        extend = []
        for job in ready:
            job.extend = False
            if self._advance(job, fetched.get(job.req.rid)):
                job.req.t_done = time.perf_counter() - t0
                done.append(job.req)
                self.used -= job.bytes
                job.release()
                del self.jobs[job.req.rid]
            elif job.extend:
                extend.append(job)
        if extend:
            self.batch.set_depths([j.slot for j in extend], [j.gs_depth for j in extend])

    def _start(self, job: _Job) -> bool:
        """Put a prepared job into the batch. False: its slot's device memory
        is not to be had while other jobs run (the budget counts bytes, but a
        slot's Krylov basis needs one contiguous block); the caller keeps the
        job and tries again when jobs have finished. With no other job
        running the error is raised."""
        job.gs_max = min(self.gs_max, job.basis.n)

        def add():
            job.slot = self.batch.add(job.hop, job.v0, job.gs_depth, max_depth=job.gs_max,
                                      keep_basis=self.keep_basis)

        try:
            add()
        except QSError as e:
            if "QS_ENOMEM_DEVICE" not in str(e):
                raise
            try:
                if self.patterns is None or not self.patterns.entries:
                    raise e
                # the cached streams may split the free device memory into
                # pieces too small for this slot's Krylov basis: drop the cache
                # (streams still in use stay with their requests), try again
                import gc
                self.patterns.clear()
                gc.collect()
                add()
            except QSError as e2:
                if "QS_ENOMEM_DEVICE" not in str(e2) or not self.jobs:
                    raise
                return False
        if self.keep_basis:
            job.kept_bytes = 16 * (job.gs_max + 1) * job.basis.n
        self.jobs[job.req.rid] = job
        self.used += job.bytes + job.kept_bytes
        return True

    def serve(self, requests: list[Request], realtime: bool = True, plain: bool = False,
              progress=None) -> list[Request]:
        """Run the stream; returns the completed requests. realtime: arrivals
        are honoured on the wall clock; plain: one request at a time, prepared
        when its turn comes (the baseline). Otherwise admitted requests are
        prepared on a worker thread while the batch steps. A request whose
        device memory alone exceeds the HBM budget is refused on arrival
        (Request.rejected says why; self.rejected lists them). progress:
        called with (completed, total, seconds) each time another 5% of the
        stream has completed."""
        pending = sorted(requests)
        waiting: list[Request] = []
        done: list[Request] = []
        prep = _Prep(self.ctx.device, self.gs_depth, self.prep_workers,
                     (self.start_temperature, self.start_floor), self.patterns, self.product_layout) \
            if (self.prepare_ahead and not plain) else None
        inflight: dict[int, int] = {}  # rid -> reserved bytes of requests being prepared
        deferred: list[_Job] = []  # prepared jobs waiting for device memory (their bytes counted in used)

        def start(job):
            if not self._start(job):
                deferred.append(job)
                self.used += job.bytes
        kept = self.gs_max + 1 if self.keep_basis else 0
        slots = 1 if plain else self.max_slots
        # The preparation threads run Python between their library calls; the
        # serving loop, which releases the interpreter lock in every library
        # call, would otherwise wait up to the default 5 ms switch interval to
        # get it back after each call while the batch on the GPU runs dry.
        switch = sys.getswitchinterval()
        if prep is not None:
            sys.setswitchinterval(2e-4)
        t0 = time.perf_counter()
        i = 0
        try:
            while i < len(pending) or waiting or self.jobs or inflight:
                now = time.perf_counter() - t0
                while i < len(pending) and (not realtime or pending[i].arrival <= now):
                    waiting.append(pending[i])
                    i += 1
                # admission: FIFO, by free slots and HBM budget; a request that
                # would not fit the budget even alone is refused at once
                while waiting and len(self.jobs) + len(inflight) + len(deferred) < slots:
                    r = waiting[0]
                    est = estimate_bytes(r.dim, kept)
                    if est > self.budget:
                        waiting.pop(0)
                        r.t_admit = r.t_done = time.perf_counter() - t0
                        r.rejected = f"needs about {est / 1e9:.1f} GB of device memory; the budget is " \
                                     f"{self.budget / 1e9:.1f} GB"
                        self.rejected.append(r)
                        continue
                    cached = self.patterns.bytes if self.patterns is not None else 0
                    if (self.jobs or inflight) and self.used + cached + est > self.budget:
                        break
                    waiting.pop(0)
                    r.t_admit = time.perf_counter() - t0
                    if prep is None:
                        start(_Job(r, self.ctx, self.gs_depth, self.start_temperature, self.start_floor,
                                   patterns=self.patterns, product=self.product_layout))
                    else:
                        inflight[r.rid] = est
                        self.used += est
                        prep.todo.put(r)
                # prepared jobs join the batch; wait for one when nothing else can run
                while prep is not None and inflight:
                    try:
                        job = prep.ready.get(block=not self.jobs, timeout=None if not self.jobs else 0)
                    except queue.Empty:
                        break
                    if isinstance(job, BaseException):
                        raise job
                    self.used -= inflight.pop(job.req.rid)
                    start(job)
                if not self.jobs:
                    if i < len(pending) and realtime and not inflight:
                        time.sleep(max(0.0, pending[i].arrival - (time.perf_counter() - t0)))
                    continue
                self.batch.run(self.chunk)
                before = len(done)
                self._step_jobs(t0, done)
                # jobs that waited for device memory, in arrival order, now
                # that finished jobs have released theirs
                while deferred:
                    self.used -= deferred[0].bytes
                    if not self._start(deferred[0]):
                        self.used += deferred[0].bytes
                        break
                    deferred.pop(0)
                if progress is not None:
                    step = max(1, len(pending) // 20)
                    if len(done) // step > before // step:
                        progress(len(done), len(pending), time.perf_counter() - t0)
        finally:
            if prep is not None:
                prep.close()
            sys.setswitchinterval(switch)
        return done


def latency_stats(done: list[Request]) -> dict:
    lat = np.array([r.t_done - r.arrival for r in done])
    span = max(r.t_done for r in done) - min(r.arrival for r in done)
    return {"n": len(done), "spectra_per_s": len(done) / span, "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()), "span_s": span}


if __name__ == "__main__":
    import sys
    if "--write-family" in sys.argv:
        with open(_FAMILY_FILE, "w") as f:
            json.dump(_count_family(), f)
