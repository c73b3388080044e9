"""Batches of independent small models (the wt_batch_* C ABI).

The reference runs DISPERSION, ASMBLE and STRUMPACK_SERIAL_FINTER jobs 2/3
once per model and frequency (elwave.f90:147-186).  A ``Batch`` does that for
many models at once, each with its own grid, medium and frequency: the
models are one block-diagonal system whose elimination forest is factored
level by level on the GPU, with the factorization and each refinement step
captured in HIP graphs.
"""
import ctypes
import json
import time

import numpy as np

from ._lib import lib, check, grid_params, WtGridParams, WtSolveStats, NotConverged, WT_ERR_NOCONV
from . import media

_vp = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(_vp)


class Batch:
    """dims: sequence of (nx, ny, nz).  max_group_nnz > 0 caps the CSR
    entries of one group (a block-diagonal system on its own stream);
    max_group_bytes > 0 caps a group's numeric state (see ``run``)."""

    def __init__(self, dims, leaf_nodes=64, max_group_nnz=0, max_group_bytes=0):
        d = np.ascontiguousarray(np.asarray(dims, np.int32).reshape(-1, 3))
        self.dims = [tuple(int(v) for v in r) for r in d]
        self.m = len(self.dims)
        h = ctypes.c_void_p()
        check(lib().wt_batch_create(self.m, _p(d), int(leaf_nodes), int(max_group_nnz), int(max_group_bytes),
                                    ctypes.byref(h)), "wt_batch_create")
        self._h = h
        self.offsets = np.zeros(self.m + 1, np.int64)
        self.group_of = np.zeros(self.m, np.int32)
        check(lib().wt_batch_offsets(h, _p(self.offsets), _p(self.group_of)), "wt_batch_offsets")
        self.n = int(self.offsets[-1])

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().wt_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        v = (ctypes.c_int64 * 8)()
        fl = ctypes.c_double()
        check(lib().wt_batch_info(self._h, v, ctypes.byref(fl)), "wt_batch_info")
        keys = ("n", "nnz", "fronts", "levels", "max_p", "groups", "factor_len", "cb_len")
        out = {k: int(v[i]) for i, k in enumerate(keys)}
        out["flops"] = fl.value
        return out

    def set_graphs(self, enable=True):
        check(lib().wt_batch_set_graphs(self._h, int(bool(enable))), "wt_batch_set_graphs")

    def set_models(self, lams, mus, rhos):
        """Per-model (lambda, mu) complex and rho real arrays, z fastest."""
        lam = np.ascontiguousarray(np.concatenate([np.asarray(a, np.complex128).ravel() for a in lams]))
        mu = np.ascontiguousarray(np.concatenate([np.asarray(a, np.complex128).ravel() for a in mus]))
        rho = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float64).ravel() for a in rhos]))
        if lam.size * 3 != self.n or mu.size != lam.size or rho.size != lam.size:
            raise ValueError("model arrays do not match the batch's grids")
        check(lib().wt_batch_set_models(self._h, _p(lam), _p(mu), _p(rho), 0), "wt_batch_set_models")

    def assemble(self, params):
        arr = (WtGridParams * self.m)(*params)
        check(lib().wt_batch_assemble(self._h, arr), "wt_batch_assemble")

    def factor(self, prec=32):
        check(lib().wt_batch_factor(self._h, prec), "wt_batch_factor")

    def autotune(self, prec=32):
        sp = ctypes.c_double()
        check(lib().wt_batch_autotune(self._h, prec, ctypes.byref(sp)), "wt_batch_autotune")
        return sp.value

    def _pack(self, blocks):
        cols = [1 if np.asarray(b).ndim == 1 else np.asarray(b).shape[1] for b in blocks]
        nrhs = max(cols)
        B = np.zeros((self.n, nrhs), np.complex128)
        for k, b in enumerate(blocks):
            b = np.asarray(b, np.complex128).reshape(self.offsets[k + 1] - self.offsets[k], -1)
            B[self.offsets[k]:self.offsets[k + 1], :b.shape[1]] = b
        return B, np.asarray(cols, np.int32), nrhs

    def _unpack(self, X, cols, shapes_1d):
        out = []
        for k in range(self.m):
            x = X[self.offsets[k]:self.offsets[k + 1], :cols[k]].copy()
            out.append(x[:, 0] if shapes_1d[k] else x)
        return out

    def apply(self, xs):
        """A_k x_k for every model (x_k: (n_k,) or (n_k, c))."""
        X, cols, nrhs = self._pack(xs)
        Y = np.zeros_like(X)
        check(lib().wt_batch_apply(self._h, nrhs, _p(X), _p(Y), 0), "wt_batch_apply")
        return self._unpack(Y, cols, [np.asarray(x).ndim == 1 for x in xs])

    def solve(self, bs, prec=32, tol=1e-10, max_iter=30):
        """Solves A_k x_k = b_k for every model; returns (xs, stats)."""
        B, cols, nrhs = self._pack(bs)
        X = np.zeros_like(B)
        st = (WtSolveStats * self.m)()
        rc = lib().wt_batch_solve(self._h, prec, nrhs, _p(cols), _p(B), _p(X), 0, float(tol), int(max_iter), st)
        stats = [{k: getattr(s, k) for k, _ in WtSolveStats._fields_} for s in st]
        xs = self._unpack(X, cols, [np.asarray(b).ndim == 1 for b in bs])
        if rc == WT_ERR_NOCONV:
            raise NotConverged("wt_batch_solve", stats, xs)
        check(rc, "wt_batch_solve")
        return xs, stats

    def run(self, bs, prec=32, tol=1e-10, max_iter=30):
        """Factor and solve group after group, releasing each group's factors
        before the next (``solve``'s arguments and results)."""
        B, cols, nrhs = self._pack(bs)
        X = np.zeros_like(B)
        st = (WtSolveStats * self.m)()
        rc = lib().wt_batch_run(self._h, prec, nrhs, _p(cols), _p(B), _p(X), 0, float(tol), int(max_iter), st)
        stats = [{k: getattr(s, k) for k, _ in WtSolveStats._fields_} for s in st]
        xs = self._unpack(X, cols, [np.asarray(b).ndim == 1 for b in bs])
        if rc == WT_ERR_NOCONV:
            raise NotConverged("wt_batch_run", stats, xs)
        check(rc, "wt_batch_run")
        return xs, stats


def symbolic_info(dims, leaf_nodes=64):
    """Host-only symbolic analysis of a batch's union (no GPU): dict as
    ``Batch.info`` without the group count."""
    d = np.ascontiguousarray(np.asarray(dims, np.int32).reshape(-1, 3))
    v = (ctypes.c_int64 * 8)()
    fl = ctypes.c_double()
    check(lib().wt_batch_symbolic_info(d.shape[0], _p(d), int(leaf_nodes), v, ctypes.byref(fl)),
          "wt_batch_symbolic_info")
    keys = ("n", "nnz", "fronts", "levels", "max_p", "factor_len", "cb_len", "w_len")
    out = {k: int(v[i]) for i, k in enumerate(keys)}
    out["flops"] = fl.value
    return out


def small_models(nmodels=256, seed=1004, side=(16, 40), freq=(0.05, 0.5), nrhs=(1, 16), dx=250.0,
                 pml_nodes=5):
    """The batched-small-models workload: every side of every grid uniform
    in ``side`` nodes, a seeded layered medium per model (as ``media.layered``),
    a frequency uniform in ``freq`` Hz and a right-hand-side count uniform in
    ``nrhs``.  Returns a list of dicts (dims, vp, vs, rho, freq, nrhs,
    params); the PML is ``pml_nodes`` intervals wide on every side but the
    free surface."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nmodels):
        nx, ny, nz = (int(v) for v in rng.integers(side[0], side[1] + 1, 3))
        f = float(rng.uniform(freq[0], freq[1]))
        c = int(rng.integers(nrhs[0], nrhs[1] + 1))
        vp1, vs1, rho1 = media.layered(nx, ny, nz, dx, seed * 1000 + k)
        shape = (ny, nx, nz)
        vp = np.broadcast_to(vp1[None, None, :], shape)
        vs = np.broadcast_to(vs1[None, None, :], shape)
        rho = np.broadcast_to(rho1[None, None, :], shape)
        p = grid_params(nx, ny, nz, dx, 2 * np.pi * f, pmld=pml_nodes * dx)
        out.append(dict(dims=(nx, ny, nz), vp=vp, vs=vs, rho=rho, freq=f, nrhs=c, params=p, dx=dx,
                        seed=seed * 1000 + k))
    return out


def plane_wave_block(model, seed_offset=0):
    """Plane-wave displacement fields (n_k, nrhs_k) of a small model at its
    frequency (known solutions: the right-hand sides are A_k U_k)."""
    nx, ny, nz = model["dims"]
    vpm = float(np.median(model["vp"]))
    return media.plane_wave_fields(nx, ny, nz, model["dx"], 2 * np.pi * model["freq"], vpm, model["nrhs"],
                                   model["seed"] + seed_offset)


def benchmark(nmodels=256, seed=1004, group_bytes=160e9, reps=2, autotune_models=48, progress=None):
    """The batched-small-models workload: ``nmodels`` (256) models, every side uniform in
    16-40 nodes, a seeded layered medium each (dx = 250 m, 5-node PML), a
    frequency uniform in 0.05-0.5 Hz and 1-16 plane-wave right-hand sides
    each (b = A u for known plane waves u).  Timed: complex64 factorization +
    FP64 refinement to 1e-8 of all models, group after group (groups bounded
    by the symbolic memory estimate), right-hand sides and solutions resident
    in HBM.  Also: the batched assembly of all 256 operators, and the
    autotuner's effect on repeated factorizations of a resident 48-model
    batch (first factorization default tiles, then tuned).  Returns a dict."""
    L = lib()
    say = progress or (lambda msg: None)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    ms = small_models(nmodels, seed=seed)
    lams, mus, rhos = [], [], []
    for m in ms:
        lam, mu, rho = media.lame(m["vp"], m["vs"], m["rho"])
        lams.append(lam)
        mus.append(mu)
        rhos.append(rho)
    t0 = time.perf_counter()
    b = Batch([m["dims"] for m in ms], max_group_bytes=int(group_bytes))
    info = b.info()
    b.set_models(lams, mus, rhos)
    setup = time.perf_counter() - t0
    params = [m["params"] for m in ms]
    b.assemble(params)
    L.wt_dev_sync()
    t0 = time.perf_counter()
    b.assemble(params)
    L.wt_dev_sync()
    t_asm = time.perf_counter() - t0
    nrhs = 16
    cols = np.array([m["nrhs"] for m in ms], np.int32)
    U = np.zeros((b.n, nrhs), np.complex128)
    for k, m in enumerate(ms):
        U[b.offsets[k]:b.offsets[k + 1], :m["nrhs"]] = plane_wave_block(m)
    dU, dB, dX = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    for d in (dU, dB, dX):
        check(L.wt_dev_alloc(U.nbytes, ctypes.byref(d)), "alloc")
    check(L.wt_h2d(dU, P(U), U.nbytes), "h2d")
    check(L.wt_batch_apply(b._h, nrhs, dU, dB, 1), "apply")
    st = (WtSolveStats * len(ms))()
    say("batched models: warm-up run")
    check(L.wt_batch_run(b._h, 32, nrhs, P(cols), dB, dX, 1, 1e-8, 30, st), "wt_batch_run")
    times = []
    for _ in range(reps):
        L.wt_dev_sync()
        t0 = time.perf_counter()
        check(L.wt_batch_run(b._h, 32, nrhs, P(cols), dB, dX, 1, 1e-8, 30, st), "wt_batch_run")
        L.wt_dev_sync()
        times.append(time.perf_counter() - t0)
    t = min(times)
    X = np.zeros_like(U)
    L.wt_d2h(P(X), dX, X.nbytes)
    werr = 0.0
    for k, m in enumerate(ms):
        a, e = b.offsets[k], b.offsets[k + 1]
        u = U[a:e, :m["nrhs"]]
        werr = max(werr, float(np.linalg.norm(X[a:e, :m["nrhs"]] - u) / np.linalg.norm(u)))
    for d in (dU, dB, dX):
        L.wt_dev_free(d)
    del U, X
    nnz = info["nnz"]
    out = {"workload": "%d layered models, sides U{16..40}, 0.05-0.5 Hz, 1-16 plane-wave RHS each, complex64 "
                       "factor + FP64 refinement to 1e-8, groups run one after the other" % len(ms),
           "models": len(ms), "rhs": int(cols.sum()), "unknowns": info["n"], "nnz": nnz,
           "fronts": info["fronts"], "groups": info["groups"], "factor_flops": info["flops"],
           "s_per_run": t, "models_per_s": len(ms) / t, "rhs_per_s": float(cols.sum()) / t,
           "factor_tflops_incl_solve": info["flops"] / t / 1e12,
           "max_rel_residual": max(x.rel_residual for x in st), "refinement_solves": max(x.iterations for x in st),
           "max_rel_wavefield_err": werr, "setup_s": setup,
           "assembly": {"ms": t_asm * 1e3, "gbs": nnz * 16 / t_asm / 1e9, "launches": len(ms)}}
    b.close()
    # autotuner on a resident batch of the first models
    sub = ms[:autotune_models]
    if not sub:
        return out
    with Batch([m["dims"] for m in sub]) as r:
        r.set_models(lams[:len(sub)], mus[:len(sub)], rhos[:len(sub)])
        r.assemble(params[:len(sub)])
        r.factor(32)
        L.wt_dev_sync()
        t0 = time.perf_counter()
        r.factor(32)
        L.wt_dev_sync()
        t_def = time.perf_counter() - t0
        sp = r.autotune(32)
        r.factor(32)
        L.wt_dev_sync()
        t0 = time.perf_counter()
        r.factor(32)
        L.wt_dev_sync()
        t_tuned = time.perf_counter() - t0
        out["autotune"] = {"models": len(sub), "factor_s_default_tiles": t_def, "factor_s_tuned_tiles": t_tuned,
                           "gemm_speedup": sp, "factor_speedup": t_def / t_tuned,
                           "note": "graph-replayed refactorization of a resident batch"}
    return out


if __name__ == "__main__":
    import argparse
    import sys
    ap = argparse.ArgumentParser(description="run the batched small-model workload once and print JSON")
    ap.add_argument("--models", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1004)
    ap.add_argument("--group-gb", type=float, default=160.0)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--autotune-models", type=int, default=48)
    a = ap.parse_args()
    print(json.dumps(benchmark(a.models, a.seed, a.group_gb * 1e9, a.reps, a.autotune_models,
                               progress=lambda m: print(m, file=sys.stderr, flush=True))))
