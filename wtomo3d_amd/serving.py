"""Serving loop for forward solves: (model, frequency, source-set) requests.

Requests are admitted as slots and HBM free up (admission by the host-side
memory estimate of the symbolic analysis, ``wt_estimate_bytes``); each slot
drives its own solver handle, so the work of concurrent requests overlaps on
the GPU through the handles' streams.  Factors stay resident in HBM after a
request and are reused by later requests for the same model and frequency;
idle factors are evicted least-recently-used when memory is needed.  The
reference has no serving layer (its driver factors one frequency at a time,
elwave.f90:147-182); this is the request-level wrapper around jobs 2 and 3.

Each request solves the reference's forward problem for plane-wave sources
on its model: a heterogeneous model (seeded by the model key) over a layered
background; Thomson-Haskell profiles of the background (HASKGRN) expanded to
3-D background fields; the scattered-field source (A_bg - A) u_bg
(HETFLDSRC); the FP32-factor / FP64-refined solve for the scattered field;
and the total field u_h + u_bg (MAKEUH).  Sources: P (70 %) or S plane
waves, incidence 10-35 degrees, azimuth 0-360 degrees, from the request's
seed.
"""
import collections
import ctypes
import threading
import time

import numpy as np

from . import media, rhs
from ._lib import lib, check, WtSolveStats, grid_params
from .solver import Solver


class Request:
    def __init__(self, rid, model_key, dims, dx, freq, nrhs, src_seed, arrival=0.0):
        self.rid, self.model_key, self.dims, self.dx = rid, model_key, tuple(dims), dx
        self.freq, self.nrhs, self.src_seed, self.arrival = freq, nrhs, src_seed, arrival
        self.start = self.end = None
        self.stats = None
        self.reused = False
        self.x = None
        self.error = None


class _Entry:
    def __init__(self, solver, nbytes):
        self.solver, self.nbytes, self.busy, self.last = solver, nbytes, False, time.time()


class Engine:
    def __init__(self, hbm_budget=200e9, slots=4, tol=1e-8, max_iter=30, keep_factors=True,
                 return_solutions=False, lookahead=8):
        self.budget, self.slots, self.tol, self.max_iter = hbm_budget, slots, tol, max_iter
        self.keep, self.return_x, self.lookahead = keep_factors, return_solutions, lookahead
        self.used = 0
        self.cache = collections.OrderedDict()  # (model_key, freq) -> _Entry
        self.pending = collections.deque()
        self.cv = threading.Condition()
        self.closed = False
        self.done = []
        self._models = {}
        self._mlock = threading.Lock()
        self._est = {}

    # -- model / memory helpers --------------------------------------------
    def _model(self, key, dims, dx):
        """(lam, mu, rho) of the heterogeneous model, its layer stack and the
        (lam, mu, rho) of its layered background."""
        k = (key, dims, dx)
        with self._mlock:
            if k not in self._models:
                nx, ny, nz = dims
                vp, vs, rho = media.heterogeneous(nx, ny, nz, dx, seed=int(key))
                vp1, vs1, rho1 = media.layered(nx, ny, nz, dx, seed=int(key))
                bg = media.lame(*(np.broadcast_to(a[None, None, :], (ny, nx, nz)) for a in (vp1, vs1, rho1)))
                self._models[k] = (media.lame(vp, vs, rho), media.layer_stack(nx, ny, nz, dx, seed=int(key)), bg)
            return self._models[k]

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

    # -- scheduling ----------------------------------------------------------
    def submit(self, req):
        with self.cv:
            self.pending.append(req)
            self.cv.notify_all()

    def _admit(self):
        """Pick a pending request that can run now (caller holds the lock)."""
        for idx in range(min(self.lookahead, len(self.pending))):
            r = self.pending[idx]
            key = (r.model_key, r.dims, r.freq)
            e = self.cache.get(key)
            if e is not None:
                if e.busy:
                    continue
                e.busy = True
                self.cache.move_to_end(key)
                del self.pending[idx]
                return r, e, 0
            need = self._estimate(r.dims)
            if need > self.budget:
                # can never fit: fail it instead of waiting forever
                del self.pending[idx]
                r.error = "needs %d bytes of HBM, budget is %d" % (need, self.budget)
                r.start = r.end = time.perf_counter()
                self.done.append(r)
                self.cv.notify_all()
                return self._admit()
            while self.used + need > self.budget:
                victim = next((k for k, v in self.cache.items() if not v.busy), None)
                if victim is None:
                    break
                v = self.cache.pop(victim)
                v.solver.close()
                self.used -= v.nbytes
            if self.used + need <= self.budget:
                self.used += need
                del self.pending[idx]
                return r, None, need
        return None, None, 0

    def _worker(self, slot_bufs):
        L = lib()
        while True:
            with self.cv:
                while True:
                    r, e, need = self._admit()
                    if r is not None or (self.closed and not self.pending):
                        break
                    self.cv.wait()
                if r is None:
                    return
            r.start = time.perf_counter()
            try:
                (lam, mu, rho), layers, bg = self._model(r.model_key, r.dims, r.dx)
                if e is None:
                    s = Solver(*r.dims)
                    s.set_model(lam, mu, rho)
                    s.assemble(grid_params(r.dims[0], r.dims[1], r.dims[2], r.dx, 2 * np.pi * r.freq,
                                           pmld=10 * r.dx))
                    s.factor(32)
                    e = _Entry(s, need)
                    e.busy = True
                else:
                    r.reused = True
                s = e.solver
                n = s.n
                st_ = s.stream()
                x, sf, ub, pu, pw, pk = slot_bufs(n * r.nrhs * 16, r.dims[2] * r.nrhs * 16, r.nrhs * 8)
                omega = 2 * np.pi * r.freq
                ain, phi, is_s = self.sources(r)
                prof = rhs.profiles(layers, ain, is_s, omega, r.dims[2], r.dx, 0.0, rhs.WT_PREC_F64, stream=st_,
                                    out=(pu, pw, pk))
                rhs.background_fields(r.dims[0], r.dims[1], r.dims[2], r.dx, prof, phi, stream=st_, out=ub)
                s.scattered_source(bg[0], bg[1], bg[2], ub, out=sf, nrhs=r.nrhs)
                st = WtSolveStats()
                check(L.wt_solver_solve(s._h, 32, r.nrhs, sf.ptr, x.ptr, 1, self.tol, self.max_iter,
                                        ctypes.byref(st)), "solve")
                rhs.total_field(x, ub, n * r.nrhs, stream=st_)
                r.stats = {k: getattr(st, k) for k, _ in WtSolveStats._fields_}
                if self.return_x:
                    r.x = x.to_host(np.complex128, (n, r.nrhs))
            except Exception as ex:  # reported per request
                r.error = str(ex)
            r.end = time.perf_counter()
            with self.cv:
                key = (r.model_key, r.dims, r.freq)
                if e is not None:
                    e.busy = False
                    e.last = time.time()
                    if self.keep and r.error is None:
                        self.cache[key] = e
                        self.cache.move_to_end(key)
                    else:
                        e.solver.close()
                        self.used -= e.nbytes
                elif need:
                    self.used -= need
                self.done.append(r)
                self.cv.notify_all()

    def run(self, requests, realtime=True):
        """Feed requests at their arrival times (or all at once) and wait."""
        L = lib()
        bufs = []

        def make_slot():
            # per-slot device buffers that only grow: solution, source and
            # background blocks, then the two profiles and the wavenumbers
            held = [None] * 6

            def get(block, prof, kh):
                sizes = (block, block, block, prof, prof, kh)
                for i, nb in enumerate(sizes):
                    if held[i] is None or held[i].nbytes < nb:
                        if held[i] is not None:
                            held[i].free()
                        held[i] = rhs.DeviceBuffer(nb)
                return held
            bufs.append(held)
            return get

        threads = [threading.Thread(target=self._worker, args=(make_slot(),), daemon=True)
                   for _ in range(self.slots)]
        t0 = time.perf_counter()
        for r in requests:
            r.arrival = r.arrival  # relative seconds
        for t in threads:
            t.start()
        for r in sorted(requests, key=lambda q: q.arrival):
            if realtime:
                dt = t0 + r.arrival - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
            r.arrival_abs = t0 + r.arrival if realtime else t0
            self.submit(r)
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for t in threads:
            t.join()
        wall = time.perf_counter() - t0
        for held in bufs:
            for b in held:
                if b is not None:
                    b.free()
        return wall

    def close(self):
        for e in self.cache.values():
            e.solver.close()
        self.cache.clear()
        self.used = 0


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


def latency_report(reqs, wall):
    ok = [r for r in reqs if r.error is None and r.end is not None]
    lat = np.array([r.end - r.arrival_abs for r in ok]) * 1e3
    svc = np.array([r.end - r.start for r in ok]) * 1e3
    rhs = sum(r.nrhs for r in ok)
    return {"requests": len(reqs), "completed": len(ok), "errors": len(reqs) - len(ok),
            "p50_latency_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "p50_service_ms": float(np.percentile(svc, 50)) if len(svc) else None,
            "rhs_per_s": rhs / wall, "wall_s": wall,
            "factor_reuse": sum(1 for r in ok if r.reused),
            "max_rel_residual": max((r.stats["rel_residual"] for r in ok), default=None)}
