"""Batched small models on the GPU (the wt_batch_* C ABI).

Each model of a batch is checked against the single-model solver on the same
model: identical assembled operators (A x agrees to rounding), and solutions
that agree to the refinement tolerance.  The right-hand sides are A U for
known plane-wave fields U, so every solution is also checked against U (the
relative wavefield error).  Graph replays are checked against the eager
launches, and the autotuned tile shapes against the default ones."""
import ctypes

import numpy as np
import pytest

import wtomo3d_amd as W
from wtomo3d_amd import media
from wtomo3d_amd.batch import Batch
from wtomo3d_amd.solver import Solver

pytestmark = pytest.mark.gpu
P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731

DIMS = [(8, 7, 6), (6, 9, 7), (10, 6, 8), (7, 7, 7), (9, 8, 6)]
FREQS = [0.12, 0.3, 0.45, 0.2, 0.08]
NCOLS = [1, 3, 2, 4, 1]
DX = 250.0


def _setup(dims=DIMS, freqs=FREQS, seed=31):
    mods = []
    for k, (d, f) in enumerate(zip(dims, freqs)):
        nx, ny, nz = d
        vp1, vs1, rho1 = media.layered(nx, ny, nz, DX, seed + k)
        shape = (ny, nx, nz)
        vp = np.broadcast_to(vp1[None, None, :], shape)
        lam, mu, rho = media.lame(vp, np.broadcast_to(vs1[None, None, :], shape),
                                  np.broadcast_to(rho1[None, None, :], shape))
        p = W.grid_params(nx, ny, nz, DX, 2 * np.pi * f, pmld=3 * DX)
        mods.append(dict(d=d, lam=lam, mu=mu, rho=rho, p=p, f=f, vp=float(np.median(vp))))
    return mods


def _batch(mods, **kw):
    b = Batch([m["d"] for m in mods], **kw)
    b.set_models([m["lam"] for m in mods], [m["mu"] for m in mods], [m["rho"] for m in mods])
    b.assemble([m["p"] for m in mods])
    return b


def _fields(mods, ncols, seed=5):
    return [media.plane_wave_fields(*m["d"], DX, 2 * np.pi * m["f"], m["vp"], c, seed + k)
            for k, (m, c) in enumerate(zip(mods, ncols))]


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("max_group_nnz", [0, 150000])
def test_batch_matches_single_model_solves(max_group_nnz):
    mods = _setup()
    with _batch(mods, max_group_nnz=max_group_nnz) as b:
        info = b.info()
        assert info["groups"] == (1 if max_group_nnz == 0 else info["groups"])
        if max_group_nnz:
            assert info["groups"] >= 2
        U = _fields(mods, NCOLS)
        B = b.apply(U)
        b.factor(32)
        X, st = b.solve(B, prec=32, tol=1e-11, max_iter=30)
    for k, m in enumerate(mods):
        assert st[k]["rel_residual"] <= 1e-11 and st[k]["nrhs"] == NCOLS[k]
        assert _rel(X[k], U[k]) <= 1e-8, k
        with Solver(*m["d"]) as s:
            s.set_model(m["lam"], m["mu"], m["rho"])
            s.assemble(m["p"])
            Bs = s.apply(U[k])
            assert _rel(B[k], Bs) <= 1e-14, k
            s.factor(64)
            Xs, sst = s.solve(Bs, prec=64, tol=1e-13, max_iter=10)
        assert _rel(X[k], Xs) <= 1e-9, k


def test_batch_graph_replays_match_eager():
    mods = _setup(seed=41)
    with _batch(mods) as b:
        U = _fields(mods, NCOLS, seed=9)
        B = b.apply(U)
        b.set_graphs(False)
        b.factor(32)
        X0, _ = b.solve(B, prec=32, tol=1e-10, max_iter=30)
        b.set_graphs(True)
        for rep in range(3):
            b.factor(32)
            Y = b.apply(U)  # eager work between the replays
            assert all(np.array_equal(y, bb) for y, bb in zip(Y, B))
            X, st = b.solve(B, prec=32, tol=1e-10, max_iter=30)
            for k in range(len(mods)):
                assert np.array_equal(X[k], X0[k]), (rep, k)


def test_batch_autotune_keeps_the_solution():
    mods = _setup(seed=51)
    with _batch(mods) as b:
        U = _fields(mods, NCOLS, seed=13)
        B = b.apply(U)
        b.factor(32)
        X0, _ = b.solve(B, prec=32, tol=1e-11, max_iter=30)
        sp = b.autotune(32)
        assert sp > 0.5
        X1, st1 = b.solve(B, prec=32, tol=1e-11, max_iter=30)
        b.factor(32)  # captured again with the tuned tile shapes
        X2, st2 = b.solve(B, prec=32, tol=1e-11, max_iter=30)
    for k in range(len(mods)):
        assert _rel(X1[k], X0[k]) <= 1e-10 and _rel(X2[k], X0[k]) <= 1e-10
        assert st2[k]["rel_residual"] <= 1e-11


def test_batch_fp64_factor():
    mods = _setup(seed=61)
    with _batch(mods) as b:
        U = _fields(mods, NCOLS, seed=17)
        B = b.apply(U)
        b.factor(64)
        X, st = b.solve(B, prec=64, tol=1e-12, max_iter=5)
    for k in range(len(mods)):
        assert st[k]["rel_residual"] <= 1e-12 and st[k]["iterations"] <= 3
        assert _rel(X[k], U[k]) <= 1e-9


def test_batch_ignores_columns_beyond_each_count():
    mods = _setup(seed=71)
    ncol = np.array([1, 2, 1, 3, 2], np.int32)
    nrhs = 4
    with _batch(mods) as b:
        U = _fields(mods, [nrhs] * len(mods), seed=21)
        B = b.apply(U)
        b.factor(32)
        Bg = np.zeros((b.n, nrhs), np.complex128)
        for k in range(len(mods)):
            Bg[b.offsets[k]:b.offsets[k + 1]] = B[k]  # every column filled
        Xg = np.full_like(Bg, 7.0)
        st = (W._lib.WtSolveStats * len(mods))()
        rc = W.lib().wt_batch_solve(b._h, 32, nrhs, P(ncol), P(Bg), P(Xg), 0, 1e-10, 30, st)
        assert rc == 0
    for k in range(len(mods)):
        xk = Xg[b.offsets[k]:b.offsets[k + 1]]
        assert np.all(xk[:, ncol[k]:] == 0)
        assert _rel(xk[:, :ncol[k]], U[k][:, :ncol[k]]) <= 1e-8
        assert st[k].nrhs == ncol[k]


def test_batch_run_releases_groups_and_matches_solve():
    """Memory-bounded groups run one after the other (wt_batch_run): every
    model's solution equals the resident-factor solve of the same batch."""
    mods = _setup(seed=81)
    with _batch(mods) as b1:
        U = _fields(mods, NCOLS, seed=23)
        B = b1.apply(U)
        b1.factor(32)
        X1, _ = b1.solve(B, prec=32, tol=1e-11, max_iter=30)
    with _batch(mods, max_group_bytes=1) as b2:  # one model per group
        assert b2.info()["groups"] == len(mods)
        X2, st = b2.run(B, prec=32, tol=1e-11, max_iter=30)
        X3, _ = b2.run(B, prec=32, tol=1e-11, max_iter=30)  # factors rebuilt
    for k in range(len(mods)):
        assert st[k]["rel_residual"] <= 1e-11
        assert _rel(X2[k], X1[k]) <= 1e-10 and np.array_equal(X2[k], X3[k])


def test_batch_edge_models():
    """A 2x2x2 model (all Dirichlet rows), a column longer than one assembly
    segment (4x3x70) and an ordinary model in one batch: each matches its
    single-model solve."""
    dims = [(2, 2, 2), (4, 3, 70), (9, 8, 7)]
    freqs = [0.2, 0.35, 0.1]
    ncols = [1, 2, 3]
    mods = _setup(dims=dims, freqs=freqs, seed=91)
    with _batch(mods) as b:
        U = _fields(mods, ncols, seed=29)
        B = b.apply(U)
        b.factor(32)
        X, st = b.solve(B, prec=32, tol=1e-11, max_iter=30)
    for k, m in enumerate(mods):
        assert st[k]["rel_residual"] <= 1e-11, k
        with Solver(*m["d"]) as s:
            s.set_model(m["lam"], m["mu"], m["rho"])
            s.assemble(m["p"])
            assert _rel(B[k], s.apply(U[k])) <= 1e-14, k
            s.factor(64)
            Xs, _ = s.solve(B[k], prec=64, tol=1e-13, max_iter=10)
        assert _rel(X[k], Xs) <= 1e-9, k
