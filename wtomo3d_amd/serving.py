"""Serving loop for forward solves: (model, frequency, source-set) requests.

Requests are admitted as slots and HBM free up (admission by the host-side
memory estimate of the symbolic analysis, ``wt_estimate_bytes``); each slot
drives its own solver handle, so the work of concurrent requests overlaps on
the GPU through the handles' streams.  Factors stay resident in HBM after a
request and are reused by later requests for the same model and frequency;
idle factors are evicted least-recently-used when memory is needed.  The
reference has no serving layer (its driver factors one frequency at a time,
elwave.f90:147-182); this is the request-level wrapper around jobs 2 and 3.

Factor cache.  A (model, frequency) key has at most one entry.  The entry is
created when its first request is admitted, before the factorization starts,
and its bytes are reserved then; a later request for the same key waits in
the queue until the entry is free and then reuses its factor (the
reference's one factorization per (model, omega), elwave.f90:171-178).  An
entry leaves the cache exactly once -- evicted while idle, dropped after a
failed request, dropped after its request when factors are not kept, or at
``close()`` -- and its bytes are released then, so ``used`` always equals
the bytes of the entries in the cache.

Shared launches.  Pending requests for new small models (at most
``group_nodes`` grid nodes) are admitted together, up to ``group_max`` of
them: their models are entered into one ``batch.Batch``, factored level by
level in the same launches, and their right-hand sides solved in one
refined block solve.  The batch stays cached under every member's key, so a
later request for any of them reuses it.

Each request solves the reference's forward problem for plane-wave sources
on its model: a heterogeneous model over a layered background; Thomson-
Haskell profiles of the background (HASKGRN) expanded to 3-D background
fields; the scattered-field source (A_bg - A) u_bg (HETFLDSRC); the
FP32-factor / FP64-refined solve for the scattered field; and the total field
u_h + u_bg (MAKEUH).  Sources: P (70 %) or S plane waves, incidence 10-35
degrees, azimuth 0-360 degrees, from the request's seed.  A request whose
refinement does not reach the engine's tolerance is reported as an error.

Models are either synthetic (seeded by the request's model key: layered
background plus a 5 % Gaussian random field) or an on-disk project, the
reference's ``name.ini`` + ``name_model.h5`` pair (READINI_FORWARD
readini.f90:34-222, MODELIO_READ_MODEL modelio.f90:45-113): its grid,
spacing, PML, free-surface and anti-alias settings come from the ini file,
the speeds and density from the HDF5 file, and its layered background is the
model's lateral average (``media.background_1d``).
"""
import collections
import ctypes
import sys
import threading
import time

import numpy as np

from . import media, rhs
from ._lib import lib, check, WtSolveStats, grid_params, NotConverged, WT_ERR_NOCONV
from .solver import Solver

_vp = ctypes.c_void_p


class Request:
    """One solve request.  ``project`` (a path prefix naming ``name.ini`` and
    ``name_model.h5``) replaces the synthetic model of ``model_key``; its
    grid and spacing are then read from the project when it is submitted."""

    def __init__(self, rid, model_key, dims, dx, freq, nrhs, src_seed, arrival=0.0, project=None):
        self.rid, self.model_key, self.dx = rid, model_key, dx
        self.dims = tuple(dims) if dims is not None else None
        self.freq, self.nrhs, self.src_seed, self.arrival = freq, nrhs, src_seed, arrival
        self.project = project
        self.start = self.end = None
        self.arrival_abs = None
        self.stats = None
        self.reused = False
        self.batched = False  # solved in a batch shared with other models
        self.x = None
        self.error = None

    def key(self):
        """The factor-cache key: one factorization per (model, frequency)."""
        if self.project is not None:
            return ("project", self.project, float(self.freq))
        return ("model", self.model_key, self.dims, float(self.dx), float(self.freq))


class _Entry:
    """One cached factor and the bytes reserved for it: a solver handle, or
    a batch of several models (``batch``, with ``index`` mapping each
    member's key to its model and ``problems`` the members' problems)."""

    def __init__(self, nbytes, group=False):
        self.solver = None
        self.batch = None
        self.group = group
        self.index = {}
        self.problems = []
        self.nbytes = nbytes
        self.busy = True     # a request owns it (factoring or solving)
        self.ready = False   # factored
        self.last = time.time()

    def close(self):
        for h in (self.solver, self.batch):
            if h is not None:
                h.close()
        self.solver = self.batch = None


class _Problem:
    """What a request needs besides its sources: the model's (lam, mu, rho),
    the layer stack and (lam, mu, rho) of its background, and the assembly
    parameters at a frequency."""

    def __init__(self, dims, dx, model, layers, bg, pmld, pmlr=1e-3, fst=True, usemin=True, tau=999.999):
        self.dims, self.dx, self.model, self.layers, self.bg = dims, dx, model, layers, bg
        self.pmld, self.pmlr, self.fst, self.usemin, self.tau = pmld, pmlr, fst, usemin, tau

    def params(self, freq):
        om = complex(2 * np.pi * freq, 0.0)
        if self.tau != 999.999:  # anti-alias damping, elwave.f90:164
            om += 1j / self.tau
        nx, ny, nz = self.dims
        return grid_params(nx, ny, nz, self.dx, om, pmld=self.pmld, pmlr=self.pmlr, fst=self.fst,
                           usemin=self.usemin)


class Engine:
    def __init__(self, hbm_budget=200e9, slots=4, tol=1e-8, max_iter=30, keep_factors=True,
                 return_solutions=False, lookahead=8, group_max=8, group_nodes=120_000, report_every=0):
        self.budget, self.slots, self.tol, self.max_iter = hbm_budget, slots, tol, max_iter
        self.keep, self.return_x, self.lookahead = keep_factors, return_solutions, lookahead
        self.group_max, self.group_nodes = group_max, group_nodes
        self.report_every = report_every  # print a line every this many finished requests
        self.used = 0
        self.cache = collections.OrderedDict()  # Request.key() -> _Entry
        self.pending = collections.deque()
        self.cv = threading.Condition()
        self.closed = False
        self.done = []
        self.factorizations = 0
        self.groups = 0
        self.peak_used = 0
        self.backlog_at_last_arrival = None
        self._problems = {}
        self._plock = threading.Lock()
        self._est = {}

    # -- models ----------------------------------------------------------------
    def _problem(self, r):
        k = ("project", r.project) if r.project is not None else ("model", r.model_key, r.dims, r.dx)
        with self._plock:
            if k not in self._problems:
                self._problems[k] = self._load_project(r.project) if r.project is not None else \
                    self._synthetic(r.model_key, r.dims, r.dx)
            return self._problems[k]

    @staticmethod
    def _synthetic(key, dims, dx):
        nx, ny, nz = dims
        vp, vs, rho = media.heterogeneous(nx, ny, nz, dx, seed=int(key))
        vp1, vs1, rho1 = media.layered(nx, ny, nz, dx, seed=int(key))
        bg = media.lame(*(np.broadcast_to(a[None, None, :], (ny, nx, nz)) for a in (vp1, vs1, rho1)))
        return _Problem(dims, dx, media.lame(vp, vs, rho), media.layer_stack(nx, ny, nz, dx, seed=int(key)), bg,
                        pmld=10 * dx)

    @staticmethod
    def _load_project(name):
        from . import io
        p, vp, vs, rho = io.load_project(name)
        nx, ny, nz, dx = p["nx"], p["ny"], p["nz"], p["dx"]
        vp1, vs1, rho1, layers = media.background_1d(vp, vs, rho, dx)
        bg = media.lame(*(np.broadcast_to(a[None, None, :], (ny, nx, nz)) for a in (vp1, vs1, rho1)))
        return _Problem((nx, ny, nz), dx, media.lame(vp, vs, rho), layers, bg, pmld=p["pmld"], pmlr=p["pmlr"],
                        fst=p["freesurf"][0], usemin=p["usemin"], tau=p["tau"])

    def _estimate(self, dims):
        """Solver bytes plus the scattered-source scratch (the background
        operator and model) and the request's background and source blocks."""
        if dims not in self._est:
            nx, ny, nz = dims
            n, nnz = 3 * nx * ny * nz, int(lib().wt_graph_nnz(nx, ny, nz))
            self._est[dims] = int(lib().wt_estimate_bytes(nx, ny, nz, 64, 32, 64)) + nnz * 16 + n * 40 + 2 * n * 64 * 16
        return self._est[dims]

    @staticmethod
    def sources(r):
        """Plane-wave sources of a request: incidence (rad), azimuth (rad), S flags."""
        rng = np.random.default_rng(r.src_seed)
        ain = rng.uniform(np.deg2rad(10.0), np.deg2rad(35.0), r.nrhs)
        phi = rng.uniform(0.0, 2 * np.pi, r.nrhs)
        is_s = (rng.random(r.nrhs) < 0.3).astype(np.int32)
        return ain, phi, is_s

    # -- GPU work (the hooks a test can replace) --------------------------------
    def _factor(self, r, prob):
        """A solver handle holding the complex64 factor of the request's
        operator (jobs 1 and 2)."""
        s = Solver(*prob.dims)
        try:
            lam, mu, rho = prob.model
            s.set_model(lam, mu, rho)
            s.assemble(prob.params(r.freq))
            s.factor(32)
        except Exception:
            s.close()
            raise
        return s

    def _solve(self, r, s, prob, slot):
        """Sources, scattered-field source, refined solve and total field of
        one request on the factored handle ``s``; fills r.stats (and r.x)."""
        L = lib()
        n = s.n
        nx, ny, nz = prob.dims
        st_ = s.stream()
        x, sf, ub, pu, pw, pk = slot(n * r.nrhs * 16, nz * r.nrhs * 16, r.nrhs * 8)
        omega = 2 * np.pi * r.freq
        ain, phi, is_s = self.sources(r)
        prof = rhs.profiles(prob.layers, ain, is_s, omega, nz, prob.dx, 0.0, rhs.WT_PREC_F64, stream=st_,
                            out=(pu, pw, pk))
        rhs.background_fields(nx, ny, nz, prob.dx, prof, phi, stream=st_, out=ub)
        s.scattered_source(prob.bg[0], prob.bg[1], prob.bg[2], ub, out=sf, nrhs=r.nrhs)
        st = WtSolveStats()
        rc = L.wt_solver_solve(s._h, 32, r.nrhs, sf.ptr, x.ptr, 1, self.tol, self.max_iter, ctypes.byref(st))
        r.stats = {k: getattr(st, k) for k, _ in WtSolveStats._fields_}
        if rc == WT_ERR_NOCONV:
            raise NotConverged("request %d: refinement stopped at %.3e > %.1e after %d solves"
                               % (r.rid, st.rel_residual, self.tol, st.iterations), r.stats)
        check(rc, "wt_solver_solve")
        rhs.total_field(x, ub, n * r.nrhs, stream=st_)
        if self.return_x:
            r.x = x.to_host(np.complex128, (n, r.nrhs))

    def _factor_group(self, rs, probs):
        """One batch holding the complex64 factors of several requests'
        operators, factored level by level together (jobs 1 and 2 for all)."""
        from .batch import Batch
        b = Batch([p.dims for p in probs])
        try:
            b.set_models([p.model[0] for p in probs], [p.model[1] for p in probs], [p.model[2] for p in probs])
            b.assemble([p.params(r.freq) for r, p in zip(rs, probs)])
            b.factor(32)
        except Exception:
            b.close()
            raise
        return b

    def _solve_group(self, rs, e, slot):
        """Sources of every request in ``rs`` (models of the batch entry
        ``e``) written into one block, one refined batch solve (the models
        not in ``rs`` masked), then each request's total field."""
        L = lib()
        b = e.batch
        st_ = ctypes.c_void_p()
        check(L.wt_batch_stream(b._h, ctypes.byref(st_)), "wt_batch_stream")
        nmax = max(r.nrhs for r in rs)
        B, X = rhs.DeviceBuffer(b.n * nmax * 16), rhs.DeviceBuffer(b.n * nmax * 16)
        ncol = np.zeros(b.m, np.int32)
        own = []
        P = lambda a: a.ctypes.data_as(_vp)  # noqa: E731
        try:
            for r in rs:
                k = e.index[r.key()]
                prob = e.problems[k]
                nx, ny, nz = prob.dims
                n = 3 * nx * ny * nz
                ain, phi, is_s = self.sources(r)
                prof = rhs.profiles(prob.layers, ain, is_s, 2 * np.pi * r.freq, nz, prob.dx, 0.0, rhs.WT_PREC_F64,
                                    stream=st_)
                ub = rhs.background_fields(nx, ny, nz, prob.dx, prof, phi, stream=st_)
                for buf in (prof.U, prof.W, prof.kh):
                    buf.free()
                own.append((r, k, n, ub))
                pr = prob.params(r.freq)
                (lam, mu, rho), (lb, mb, rb) = prob.model, prob.bg
                dst = _vp(B.ptr.value + int(b.offsets[k]) * nmax * 16)
                check(L.wt_scattered_source_grid(ctypes.byref(pr), P(lam), P(mu), P(rho), P(lb), P(mb), P(rb), 0,
                                                 r.nrhs, ub.ptr, r.nrhs, None, dst, nmax, st_),
                      "wt_scattered_source_grid")
                ncol[k] = r.nrhs
            stats = (WtSolveStats * b.m)()
            rc = L.wt_batch_solve(b._h, 32, nmax, P(ncol), B.ptr, X.ptr, 1, self.tol, self.max_iter, stats)
            if rc != WT_ERR_NOCONV:
                check(rc, "wt_batch_solve")
            for r, k, n, ub in own:
                st = stats[k]
                r.stats = {f: getattr(st, f) for f, _ in WtSolveStats._fields_}
                if not st.converged:
                    r.error = "request %d: refinement stopped at %.3e > %.1e after %d solves" % (
                        r.rid, st.rel_residual, self.tol, st.iterations)
                    continue
                x = rhs.DeviceBuffer(n * r.nrhs * 16)
                try:
                    check(L.wt_dev_copy2d(x.ptr, r.nrhs * 16, _vp(X.ptr.value + int(b.offsets[k]) * nmax * 16),
                                          nmax * 16, r.nrhs * 16, n, st_), "wt_dev_copy2d")
                    rhs.total_field(x, ub, n * r.nrhs, stream=st_)
                    if self.return_x:
                        r.x = x.to_host(np.complex128, (n, r.nrhs))
                finally:
                    x.free()
        finally:
            for _, _, _, ub in own:
                ub.free()
            B.free()
            X.free()

    def _slot(self):
        """Per-slot device buffers that only grow: solution, source and
        background blocks, then the two profiles and the wavenumbers.
        Returns (get(block, prof, kh) -> buffers, release())."""
        held = [None] * 6

        def get(block, prof, kh):
            sizes = (block, block, block, prof, prof, kh)
            for i, nb in enumerate(sizes):
                if held[i] is None or held[i].nbytes < nb:
                    if held[i] is not None:
                        held[i].free()
                    held[i] = rhs.DeviceBuffer(nb)
            return held

        def release():
            for b in held:
                if b is not None:
                    b.free()
        return get, release

    # -- scheduling ----------------------------------------------------------
    def submit(self, req):
        if req.project is not None:
            prob = self._problem(req)
            req.dims, req.dx = prob.dims, prob.dx
        with self.cv:
            self.pending.append(req)
            self.cv.notify_all()

    def _drop(self, e):
        """Remove an entry (under every key it serves) from the cache and
        release its bytes -- once; the caller holds the lock and no request
        is using the entry."""
        for k in [k for k, v in self.cache.items() if v is e]:
            del self.cache[k]
        e.close()
        self.used -= e.nbytes
        e.nbytes = 0

    def _small(self, r):
        return self.group_max > 1 and r.dims[0] * r.dims[1] * r.dims[2] <= self.group_nodes

    def _reserve(self, need):
        """Evict idle entries, least recently used first, until ``need``
        more bytes fit; True when they do."""
        while self.used + need > self.budget:
            victim = next((v for v in self.cache.values() if not v.busy), None)
            if victim is None:
                break
            self._drop(victim)
        return self.used + need <= self.budget

    def _admit(self):
        """Pick pending requests that can run now and the cache entry they
        run on (caller holds the lock): a request on a cached factor, a
        group of new small models sharing one batch, or one new model."""
        idx = 0
        while idx < min(self.lookahead, len(self.pending)):
            r = self.pending[idx]
            key = r.key()
            e = self.cache.get(key)
            if e is not None:
                if e.busy:  # its factor is being built or used: wait for it
                    idx += 1
                    continue
                e.busy = True
                self.cache.move_to_end(key)
                del self.pending[idx]
                return [r], e
            need = self._estimate(r.dims)
            if need > self.budget:
                # can never fit: fail it instead of waiting forever
                del self.pending[idx]
                r.error = "needs %d bytes of HBM, budget is %d" % (need, self.budget)
                r.start = r.end = time.perf_counter()
                self.done.append(r)
                self.cv.notify_all()
                continue
            if self._small(r):
                group, keys = [r], {key}
                for q in list(self.pending)[idx + 1:self.lookahead]:
                    kq = q.key()
                    if kq not in keys and kq not in self.cache and self._small(q):
                        group.append(q)
                        keys.add(kq)
                        if len(group) == self.group_max:
                            break
                if len(group) > 1:
                    gneed = sum(self._estimate(q.dims) for q in group)
                    if gneed <= self.budget and self._reserve(gneed):
                        e = _Entry(gneed, group=True)
                        for q in group:
                            self.cache[q.key()] = e
                            self.pending.remove(q)
                        self.used += gneed
                        self.peak_used = max(self.peak_used, self.used)
                        return group, e
            if self._reserve(need):
                e = _Entry(need)
                self.cache[key] = e
                self.used += need
                self.peak_used = max(self.peak_used, self.used)
                del self.pending[idx]
                return [r], e
            idx += 1
        return None, None

    def _worker(self, slot):
        while True:
            with self.cv:
                while True:
                    rs, e = self._admit()
                    if rs is not None or (self.closed and not self.pending):
                        break
                    self.cv.wait()
                if rs is None:
                    return
            t0 = time.perf_counter()
            for r in rs:
                r.start = t0
                r.batched = e.group
            try:
                if e.group:
                    if not e.ready:
                        probs = [self._problem(r) for r in rs]
                        e.batch = self._factor_group(rs, probs)
                        e.problems = probs
                        e.index = {r.key(): k for k, r in enumerate(rs)}
                        e.ready = True
                        self.factorizations += len(rs)
                        self.groups += 1
                    else:
                        for r in rs:
                            r.reused = True
                    self._solve_group(rs, e, slot)
                else:
                    r = rs[0]
                    prob = self._problem(r)
                    if not e.ready:
                        e.solver = self._factor(r, prob)
                        e.ready = True
                        self.factorizations += 1
                    else:
                        r.reused = True
                    self._solve(r, e.solver, prob, slot)
            except Exception as ex:  # reported per request
                for r in rs:
                    if r.error is None:
                        r.error = str(ex)
            t1 = time.perf_counter()
            with self.cv:
                for r in rs:
                    r.end = t1
                e.busy = False
                e.last = time.time()
                if any(r.error is not None for r in rs) or not self.keep:
                    # a failed request's factor is not trusted by later ones
                    self._drop(e)
                before = len(self.done)
                self.done.extend(rs)
                if self.report_every and len(self.done) // self.report_every > before // self.report_every:
                    print("serving: %d requests finished, %d pending, %.0f GB reserved"
                          % (len(self.done), len(self.pending), self.used / 1e9), file=sys.stderr, flush=True)
                self.cv.notify_all()

    def run(self, requests, realtime=True):
        """Feed requests at their arrival times (or all at once) and wait."""
        slots = [self._slot() for _ in range(self.slots)]
        threads = [threading.Thread(target=self._worker, args=(get,), daemon=True) for get, _ in slots]
        with self.cv:
            self.closed = False
        t0 = time.perf_counter()
        for t in threads:
            t.start()
        for r in sorted(requests, key=lambda q: q.arrival):
            if realtime:
                dt = t0 + r.arrival - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
            r.arrival_abs = t0 + r.arrival if realtime else t0
            try:
                self.submit(r)
            except Exception as ex:  # e.g. an unreadable project
                r.error = str(ex)
                r.start = r.end = time.perf_counter()
                with self.cv:
                    self.done.append(r)
        with self.cv:
            # requests arrived but not yet finished when the last one arrives
            self.backlog_at_last_arrival = sum(1 for r in requests if r.end is None and r.error is None)
            self.closed = True
            self.cv.notify_all()
        for t in threads:
            t.join()
        wall = time.perf_counter() - t0
        for _, release in slots:
            release()
        return wall

    def accounted(self):
        """True when the reserved bytes equal those of the cached entries
        (an entry serving several keys counted once)."""
        with self.cv:
            return self.used == sum(e.nbytes for e in {id(v): v for v in self.cache.values()}.values())

    def close(self):
        with self.cv:
            for e in list({id(v): v for v in self.cache.values()}.values()):
                self._drop(e)
            self.cache.clear()


def poisson_stream(rate, duration, seed, side=(32, 80), freq=(0.05, 0.5), nrhs=(1, 64), reuse=0.3, dx=200.0,
                   n_models=None):
    """The serving workload: Poisson arrivals at `rate` per second for
    `duration` seconds; each request a (model, frequency, source-set) solve on
    a heterogeneous model whose sides are uniform in `side`, frequency uniform
    in `freq`, 1-64 plane-wave sources; a fraction `reuse` repeats an earlier
    (model, frequency) so its factor can be reused."""
    rng = np.random.default_rng(seed)
    t, reqs, seen = 0.0, [], []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        if seen and rng.random() < reuse:
            key, dims, f = seen[rng.integers(len(seen))]
        else:
            key = int(rng.integers(1, 1 << 30))
            dims = tuple(int(v) for v in rng.integers(side[0], side[1] + 1, 3))
            f = float(rng.uniform(freq[0], freq[1]))
            seen.append((key, dims, f))
        nr = int(rng.integers(nrhs[0], nrhs[1] + 1))
        reqs.append(Request(len(reqs), key, dims, dx, f, nr, int(rng.integers(1 << 30)), t))
    return reqs


def project_requests(name, nrhs, src_seed, start_id=0):
    """One request per frequency of an on-disk project (the frequencies of
    its ini file, readini.f90:168-178), all arriving at once."""
    from . import io
    p = io.read_ini(name + ".ini")
    return [Request(start_id + i, None, None, p["dx"], f, nrhs, src_seed + i, 0.0, project=name)
            for i, f in enumerate(p["freq"])]


def latency_report(reqs, wall):
    ok = [r for r in reqs if r.error is None and r.end is not None]
    lat = np.array([r.end - r.arrival_abs for r in ok]) * 1e3
    svc = np.array([r.end - r.start for r in ok]) * 1e3
    rhs_ = sum(r.nrhs for r in ok)
    return {"requests": len(reqs), "completed": len(ok), "errors": len(reqs) - len(ok),
            "p50_latency_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "p50_service_ms": float(np.percentile(svc, 50)) if len(svc) else None,
            "rhs_per_s": rhs_ / wall, "wall_s": wall,
            "factor_reuse": sum(1 for r in ok if r.reused),
            "batched": sum(1 for r in ok if r.batched),
            "max_rel_residual": max((r.stats["rel_residual"] for r in ok), default=None),
            "error_samples": sorted({str(r.error) for r in reqs if r.error is not None})[:5]}


def run_stream(rate, seconds, seed=1005, slots=4, hbm_budget=200e9, side=(32, 80), report_every=20, **engine_kw):
    """One realtime run of the serving workload (``poisson_stream``) at
    ``rate`` requests/s for ``seconds`` of arrivals.  Returns the latency
    report plus the backlog when the last request arrived (requests arrived
    and not finished), the p50 latency of the first and last third of the
    arrivals, the peak reserved HBM, the factorizations and the batched
    groups."""
    reqs = poisson_stream(rate, seconds, seed=seed, side=side, freq=(0.05, 0.5), nrhs=(1, 64), reuse=0.3, dx=200.0)
    eng = Engine(hbm_budget=hbm_budget, slots=slots, tol=1e-8, max_iter=30, report_every=report_every, **engine_kw)
    wall = eng.run(reqs, realtime=True)
    eng.close()
    rep = latency_report(reqs, wall)
    ok = sorted((r for r in reqs if r.error is None), key=lambda r: r.arrival)
    third = max(len(ok) // 3, 1)
    lat = lambda rs: float(np.percentile([(r.end - r.arrival_abs) * 1e3 for r in rs], 50)) if rs else None  # noqa
    rep.update(rate_per_s=rate, window_s=seconds, slots=slots, sides="%d-%d" % side,
               backlog_at_last_arrival=eng.backlog_at_last_arrival, p50_first_third_ms=lat(ok[:third]),
               p50_last_third_ms=lat(ok[-third:]), peak_reserved_gb=eng.peak_used / 1e9,
               factorizations=eng.factorizations, groups=eng.groups,
               completed_per_s=len(ok) / wall if wall > 0 else None)
    first, last = rep["p50_first_third_ms"], rep["p50_last_third_ms"]
    rep["sustained"] = bool(rep["errors"] == 0 and rep["backlog_at_last_arrival"] is not None and
                            rep["backlog_at_last_arrival"] <= 2 * slots and first and last and last <= 2.0 * first)
    return rep


def sustained_rate(rates, seconds, **kw):
    """``run_stream`` at each rate in ascending order (stopping after the
    first rate that is not sustained); returns (highest sustained rate or
    None, [reports])."""
    best, reps = None, []
    for rate in sorted(rates):
        rep = run_stream(rate, seconds, **kw)
        reps.append(rep)
        print("serving: %.2f req/s -> p50 %.0f ms, p99 %.0f ms, backlog %s, sustained %s"
              % (rate, rep["p50_latency_ms"] or -1, rep["p99_latency_ms"] or -1, rep["backlog_at_last_arrival"],
                 rep["sustained"]), file=sys.stderr, flush=True)
        if not rep["sustained"]:
            break
        best = rate
    return best, reps


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="serving workload: sustained-rate search and bounded bursts")
    ap.add_argument("--rates", default="1,2,3", help="comma-separated request rates to try, ascending")
    ap.add_argument("--seconds", type=float, default=150.0, help="arrival window per rate")
    ap.add_argument("--burst-rate", type=float, default=20.0)
    ap.add_argument("--burst-seconds", type=float, default=0.0, help="also run a bounded window at --burst-rate")
    ap.add_argument("--group-max", type=int, default=8)
    ap.add_argument("--out", default=None, help="write the reports as JSON here")
    a = ap.parse_args()
    best, reps = sustained_rate([float(v) for v in a.rates.split(",")], a.seconds, group_max=a.group_max)
    out = {"sustained_rate_per_s": best, "runs": reps}
    if a.burst_seconds > 0:
        out["burst"] = run_stream(a.burst_rate, a.burst_seconds, group_max=a.group_max)
        print("serving: burst %.1f req/s for %.1f s -> %s" % (a.burst_rate, a.burst_seconds,
                                                            json.dumps(out["burst"])), file=sys.stderr, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(out))
