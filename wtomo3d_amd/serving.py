"""Serving loop for forward solves: (model, frequency, source-set) requests.

Requests are admitted as slots and HBM free up (admission by the host-side
memory estimate of the symbolic analysis, ``wt_estimate_bytes``); each slot
drives its own solver handle, so the work of concurrent requests overlaps on
the GPU through the handles' streams.  Factors stay resident in HBM after a
request and are reused by later requests for the same model and frequency;
idle factors are evicted least-recently-used when memory is needed.  The
reference has no serving layer (its driver factors one frequency at a time,
elwave.f90:147-182); this is the request-level wrapper around jobs 2 and 3.
"""
import collections
import ctypes
import threading
import time

import numpy as np

from . import media
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
        self._est = {}

    # -- model / memory helpers --------------------------------------------
    def _model(self, key, dims, dx):
        k = (key, dims, dx)
        if k not in self._models:
            vp, vs, rho = media.heterogeneous(dims[0], dims[1], dims[2], dx, seed=int(key))
            lam, mu, rho1 = media.lame(vp, vs, rho)
            self._models[k] = (lam, mu, rho1, float(np.median(vp)))
        return self._models[k]

    def _estimate(self, dims):
        if dims not in self._est:
            self._est[dims] = int(lib().wt_estimate_bytes(dims[0], dims[1], dims[2], 64, 32, 64))
        return self._est[dims]

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
                if e is None:
                    lam, mu, rho, _ = self._model(r.model_key, r.dims, r.dx)
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
                vref = self._model(r.model_key, r.dims, r.dx)[3]
                bufs = slot_bufs(3 * n * r.nrhs * 16)
                dirs = np.ascontiguousarray(media.plane_wave_dirs(r.nrhs, r.src_seed), np.float64)
                omega = 2 * np.pi * r.freq
                check(L.wt_solver_planewaves(s._h, r.nrhs, dirs.ctypes.data_as(ctypes.c_void_p), omega / vref,
                                             r.dx, bufs[0]), "planewaves")
                check(L.wt_solver_apply(s._h, r.nrhs, bufs[0], bufs[1], 1), "apply")
                st = WtSolveStats()
                check(L.wt_solver_solve(s._h, 32, r.nrhs, bufs[1], bufs[2], 1, self.tol, self.max_iter,
                                        ctypes.byref(st)), "solve")
                r.stats = {k: getattr(st, k) for k, _ in WtSolveStats._fields_}
                if self.return_x:
                    x = np.zeros((n, r.nrhs), np.complex128)
                    check(L.wt_d2h(x.ctypes.data_as(ctypes.c_void_p), bufs[2], x.nbytes), "d2h")
                    r.x = x
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
            cap = [0, [ctypes.c_void_p() for _ in range(3)]]

            def get(nbytes):
                if nbytes > cap[0]:
                    for p in cap[1]:
                        if p.value:
                            L.wt_dev_free(p)
                    cap[1] = [ctypes.c_void_p() for _ in range(3)]
                    for p in cap[1]:
                        check(L.wt_dev_alloc(nbytes // 3, ctypes.byref(p)), "slot alloc")
                    cap[0] = nbytes
                return cap[1]
            bufs.append(cap)
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
        for cap in bufs:
            for p in cap[1]:
                if p.value:
                    L.wt_dev_free(p)
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
