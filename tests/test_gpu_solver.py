"""GPU multifrontal factor + block solve against the CPU direct-solve oracle.

The reference's own solve (STRUMPACK) is not available, so the solution is
pinned by an independent FP64 SuperLU solve of the same system and by the
FP64 residual; tolerances are stated per test."""
import ctypes
import os
import sys

import numpy as np
import pytest

import wtomo3d_amd as W
from wtomo3d_amd.rhs import DeviceBuffer
from wtomo3d_amd.solver import Solver
from solve_cases import problem, rhs_block
from helpers import rel_err_entries

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import solve_oracle as SO  # noqa: E402

pytestmark = pytest.mark.gpu

# (nx, ny, nz, leaf nodes): single-level, multi-level, multi-tile separators
CASES = [(5, 4, 6, 64), (10, 9, 11, 64), (16, 15, 14, 64), (12, 12, 12, 8), (9, 22, 7, 27)]


def _uniform_all_pml(n=14):
    """Uniform half-space where the 10-node PML covers the whole grid (the
    24^3 plain-method setting, shrunk): stresses pivoting."""
    import oracle_lib as O
    vp = np.full((n, n, n), 6000.0)
    lam, mu = O.lame(vp, vp / np.sqrt(3.0), np.full_like(vp, 2700.0), 64)
    ipml, pmlf = O.pml_params(500.0, 5000.0, 1e-3, 64)
    rc, vals = O.assemble(n, n, n, 500.0, 2 * np.pi * 0.2, lam, mu, np.full(n ** 3, 2700.0), 5000.0, pmlf, ipml,
                          prec=64)
    assert rc == 0
    ir, jc = O.graph(n, n, n)
    return lam, mu, ipml, pmlf, ir, jc, vals


def _setup(nx, ny, nz, leaf, seed=1):
    prm, lam, mu, rho, ir, jc, vals = problem(nx, ny, nz, seed=seed)
    A = SO.csr(ir, jc, vals)
    s = Solver(nx, ny, nz, leaf_nodes=leaf)
    s.set_model(lam, mu, rho)
    g = W._lib.WtGridParams(nx, ny, nz, prm["ipml"], 1, 1, prm["dx"], prm["pmld"], prm["pmlf"], prm["omega"], 0.0)
    s.assemble(g)
    return s, A, vals


@pytest.mark.parametrize("case", CASES)
def test_fp64_factor_solve_matches_superlu(case):
    nx, ny, nz, leaf = case
    s, A, vals = _setup(nx, ny, nz, leaf)
    v = s.values()  # FP64 assembly on the GPU == oracle (to rounding)
    assert np.max(np.abs(v - vals)) <= 1e-15 * np.max(np.abs(vals))
    B, Xt = rhs_block(A, 3, 11)
    Xo = SO.solve_superlu(A, B)
    s.factor(64)
    X, st = s.solve(B, prec=64, tol=0.0, max_iter=1)
    assert st["iterations"] == 1
    # backward error at FP64 level; forward error as small as the
    # independent direct solver's (the PML operator has cond ~1e8-1e9)
    assert st["rel_residual"] <= 1e-11
    assert SO.rel_residual(A, X, B) <= 1e-11
    err_oracle = np.max(np.abs(Xo - Xt)) / np.max(np.abs(Xt))
    err_gpu = np.max(np.abs(X - Xt)) / np.max(np.abs(Xt))
    assert err_gpu <= 10 * err_oracle + 1e-13
    s.close()


@pytest.mark.parametrize("case", CASES)
def test_fp32_factor_with_fp64_refinement(case):
    nx, ny, nz, leaf = case
    s, A, vals = _setup(nx, ny, nz, leaf, seed=2)
    B, Xt = rhs_block(A, 4, 12)
    Xo = SO.solve_superlu(A, B)
    s.factor(32)
    X, st = s.solve(B, prec=32, tol=1e-10, max_iter=30)
    assert st["rel_residual"] <= 1e-10
    assert SO.rel_residual(A, X, B) <= 1e-10
    err_oracle = np.max(np.abs(Xo - Xt)) / np.max(np.abs(Xt))
    err_gpu = np.max(np.abs(X - Xt)) / np.max(np.abs(Xt))
    assert err_gpu <= 10 * err_oracle + 1e-12
    assert st["rel_residual_first"] < 1e-2  # the complex64 factor alone is a good preconditioner
    s.close()


@pytest.mark.parametrize("dims", [(2, 2, 2), (2, 3, 4), (3, 2, 5)])
def test_tiny_grids_against_dense_lapack(dims):
    """The smallest grids (one or a few fronts, pivot blocks narrower than a
    64-row tile, PML over every node): complex128 factor + one solve and
    complex64 factor + refinement against numpy's dense LAPACK solve."""
    s, A, vals = _setup(*dims, 64, seed=3)
    B, Xt = rhs_block(A, 3, 17)
    Ad = A.toarray()
    Xd = np.linalg.solve(Ad, B)
    scale = np.max(np.abs(Xd))
    s.factor(64)
    X, st = s.solve(B, prec=64, tol=0.0, max_iter=1)
    assert SO.rel_residual(A, X, B) <= 1e-12
    assert np.max(np.abs(X - Xd)) <= 1e-10 * scale
    s.factor(32)
    X, st = s.solve(B, prec=32, tol=1e-12, max_iter=30)
    assert st["converged"] == 1 and SO.rel_residual(A, X, B) <= 1e-12
    assert np.max(np.abs(X - Xd)) <= 1e-10 * scale
    s.close()


def test_block_solve_equals_column_solves():
    """Each right-hand side is solved the same alone and inside a block."""
    s, A, vals = _setup(10, 9, 11, 64)
    B, _ = rhs_block(A, 5, 13)
    s.factor(32)
    Xb, _ = s.solve(B, prec=32, tol=0.0, max_iter=3)
    for j in range(5):
        xj, _ = s.solve(B[:, j], prec=32, tol=0.0, max_iter=3)
        assert np.array_equal(xj, Xb[:, j])
    s.close()


def test_refactor_same_handle_new_frequency():
    nx, ny, nz = 10, 9, 11
    prm, lam, mu, rho, ir, jc, vals = problem(nx, ny, nz, seed=4)
    s = Solver(nx, ny, nz)
    s.set_model(lam, mu, rho)
    for f in (0.15, 0.4):
        om = 2 * np.pi * f
        g = W._lib.WtGridParams(nx, ny, nz, prm["ipml"], 1, 1, prm["dx"], prm["pmld"], prm["pmlf"], om, 0.0)
        s.assemble(g)
        A = SO.csr(ir, jc, s.values())
        B, _ = rhs_block(A, 2, 14)
        s.factor(64)
        X, st = s.solve(B, prec=64, tol=0.0, max_iter=1)
        assert SO.rel_residual(A, X, B) <= 1e-11
    s.close()


def test_autotuned_tiles_keep_the_solution():
    """The tile-shape autotuner (one factorization per GEMM tile variant)
    leaves the handle factored, and the tuned factorization solves to the
    same refined solution."""
    nx, ny, nz = 16, 14, 12
    prm, lam, mu, rho, ir, jc, vals = problem(nx, ny, nz, seed=8)
    s = Solver(nx, ny, nz)
    s.set_values(vals)
    A = SO.csr(ir, jc, s.values())
    B, _ = rhs_block(A, 3, 21)
    s.factor(32)
    X0, _ = s.solve(B, prec=32, tol=1e-12, max_iter=30)
    sp = s.autotune(32)
    assert sp > 0.5
    X1, st1 = s.solve(B, prec=32, tol=1e-12, max_iter=30)
    s.factor(32)
    X2, st2 = s.solve(B, prec=32, tol=1e-12, max_iter=30)
    s.close()
    for X, st in ((X1, st1), (X2, st2)):
        assert st["rel_residual"] <= 1e-12
        assert np.linalg.norm(X - X0) / np.linalg.norm(X0) <= 1e-10


def test_tile_table_round_trip():
    """Stored tile shapes (wt_solver_get_tiles / wt_solver_set_tiles) apply
    to a fresh handle of the same grid and give the tuned handle's solution;
    a table of the wrong length is refused."""
    nx, ny, nz = 14, 12, 10
    prm, lam, mu, rho, ir, jc, vals = problem(nx, ny, nz, seed=9)
    A = None
    with Solver(nx, ny, nz) as s:
        s.set_values(vals)
        A = SO.csr(ir, jc, s.values())
        B, _ = rhs_block(A, 3, 31)
        s.autotune(32)
        X1, _ = s.solve(B, prec=32, tol=1e-12, max_iter=30)
        tf, ts = s.get_tiles(32, 0), s.get_tiles(32, 1, 3)
    # complex64 plans may use the four fp32 tiles and the two bf16-split tiles
    assert tf.size > 0 and ts.size > 0 and set(np.unique(np.concatenate([tf, ts]))) <= set(range(6))
    with Solver(nx, ny, nz) as s:
        s.set_values(vals)
        s.set_tiles(tf, 32, 0)
        with pytest.raises(RuntimeError):
            s.set_tiles(tf[:-1], 32, 0)
        s.factor(32)
        s.set_tiles(ts, 32, 1, 3)
        assert np.array_equal(s.get_tiles(32, 0), tf) and np.array_equal(s.get_tiles(32, 1, 3), ts)
        X2, st = s.solve(B, prec=32, tol=1e-12, max_iter=30)
    assert st["rel_residual"] <= 1e-12
    assert np.linalg.norm(X2 - X1) / np.linalg.norm(X1) <= 1e-10


def test_strumpack_serial_finter_dropin():
    """Jobs 1/2/3/-1 with the reference's argument list (strumpack_finter.c:9)."""
    nx, ny, nz = 8, 7, 9
    prm, lam, mu, rho, ir, jc, vals64 = problem(nx, ny, nz, seed=6)
    vals = vals64.astype(np.complex64)
    n = ir.size - 1
    A = SO.csr(ir, jc, vals.astype(np.complex128))
    b = np.asarray(A @ (np.ones(n) + 0.5j)).astype(np.complex64)
    x = np.zeros(n, np.complex64)
    L = W.lib()
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    ierr = ctypes.c_int(-1)
    ir0, jc0 = ir.copy(), jc.copy()
    # solve before init / factor: error, like the reference
    L.strumpack_serial_finter(I(3), I(n), I(jc.size), P(ir), P(jc), P(vals), P(b), P(x), ctypes.byref(ierr))
    assert ierr.value == 1
    L.strumpack_serial_finter(I(2), I(n), I(jc.size), P(ir), P(jc), P(vals), P(b), P(x), ctypes.byref(ierr))
    assert ierr.value == 1
    L.strumpack_serial_finter(I(1), I(n), I(jc.size), P(ir), P(jc), P(vals), P(b), P(x), ctypes.byref(ierr))
    assert ierr.value == 0
    L.strumpack_serial_finter(I(1), I(n), I(jc.size), P(ir), P(jc), P(vals), P(b), P(x), ctypes.byref(ierr))
    assert ierr.value == 1  # already initialized
    L.strumpack_serial_finter(I(2), I(n), I(jc.size), P(ir), P(jc), P(vals), P(b), P(x), ctypes.byref(ierr))
    assert ierr.value == 0
    L.strumpack_serial_finter(I(3), I(n), I(jc.size), P(ir), P(jc), P(vals), P(b), P(x), ctypes.byref(ierr))
    assert ierr.value == 0
    assert np.array_equal(ir, ir0) and np.array_equal(jc, jc0)  # pattern left untouched
    r = np.linalg.norm(b.astype(np.complex128) - A @ x.astype(np.complex128)) / np.linalg.norm(b)
    assert r <= 1e-4  # rctol of strumpack_finter.c:39, on the complex64 output
    L.strumpack_serial_finter(I(-1), I(n), I(jc.size), None, None, None, None, None, ctypes.byref(ierr))
    assert ierr.value == 0
    # a pattern that is not a 27-point grid is rejected at job 1
    bad = jc.copy()
    bad[5], bad[6] = bad[6], bad[5]
    L.strumpack_serial_finter(I(1), I(n), I(jc.size), P(ir), P(bad), P(vals), P(b), P(x), ctypes.byref(ierr))
    assert ierr.value == 1


@pytest.mark.parametrize("prec", [32, 64])
def test_all_pml_uniform_model(prec):
    n = 14
    lam, mu, ipml, pmlf, ir, jc, vals = _uniform_all_pml(n)
    A = SO.csr(ir, jc, vals)
    B, Xt = rhs_block(A, 2, 21)
    Xo = SO.solve_superlu(A, B)
    s = Solver(n, n, n)
    s.set_model(lam, mu, np.full(n ** 3, 2700.0))
    s.assemble(W._lib.WtGridParams(n, n, n, ipml, 1, 1, 500.0, 5000.0, pmlf, 2 * np.pi * 0.2, 0.0))
    s.factor(prec)
    X, st = s.solve(B, prec=prec, tol=1e-10, max_iter=30)
    assert st["rel_residual"] <= 1e-10
    err_oracle = np.max(np.abs(Xo - Xt)) / np.max(np.abs(Xt))
    assert np.max(np.abs(X - Xt)) / np.max(np.abs(Xt)) <= 10 * err_oracle + 1e-12
    s.close()


@pytest.mark.parametrize("prec", [32, 64])
def test_singular_operator_reports_zero_pivot(prec):
    """A CSR row of zeros makes the operator exactly singular: every update
    that row receives is zero, so its diagonal tile meets an exactly zero
    pivot and the factorization returns WT_ERR_PIVOT (7).  The all-zero
    operator does the same from the first tile on, and the handle factors
    normally again once good values are set."""
    nx, ny, nz = 10, 9, 11
    s, A, vals = _setup(nx, ny, nz, 64, seed=6)
    ir = A.indptr
    r = 3 * (5 + 4 * nz + 3 * nx * nz) + 1  # an interior unknown
    v = vals.copy()
    v[ir[r]:ir[r + 1]] = 0
    s.set_values(v)
    assert W.lib().wt_solver_factor(s._h, prec) == 7
    rep = s.pivot_report(prec)
    assert rep["zero_pivots"] == 1 and rep["tiny_pivots"] >= 1 and rep["min_pivot_ratio"] == 0.0
    s.set_values(np.zeros_like(vals))
    assert W.lib().wt_solver_factor(s._h, prec) == 7
    s.set_values(vals)
    s.factor(prec)
    B, Xt = rhs_block(A, 2, 31)
    X, st = s.solve(B, prec=prec, tol=1e-10, max_iter=30)
    assert st["rel_residual"] <= 1e-10
    s.close()


@pytest.mark.parametrize("prec", [32, 64])
def test_near_singular_operator_pivot_report(prec):
    """The v-row of an interior node replaced by its u-row plus 1e-10 of the
    original v-row (the two rows share their columns): near-singular, not
    singular, with the near-dependence kept through rounding to complex64
    (the smallest singular value is the rounding of the copied row, about
    1e-7 of the row, where an eigenvalue shift of this non-normal operator
    is moved far off by the same rounding).  The factorization succeeds,
    and its pivot report flags tiny pivots and a least pivot ratio far below
    the unmodified operator's, which has none.  The complex64 factor shares
    the near-dependence, so refinement of a consistent block still reaches
    1e-8; what it reports is checked against the host's FP64 residual, and
    a tolerance below its reach ends with NotConverged (converged = 0)
    carrying the residual it did reach."""
    from wtomo3d_amd._lib import NotConverged
    nx, ny, nz = 6, 5, 7
    s, A, vals = _setup(nx, ny, nz, 64, seed=9)
    s.factor(prec)
    reg = s.pivot_report(prec)
    node = 2 + 2 * nz + 3 * nx * nz  # interior: z = 3, x = 3, y = 4 (1-based)
    r1, r2 = 3 * node, 3 * node + 1
    ip = A.indptr
    assert np.array_equal(A.indices[ip[r1]:ip[r1 + 1]], A.indices[ip[r2]:ip[r2 + 1]])
    v = vals.copy()
    v[ip[r2]:ip[r2 + 1]] = vals[ip[r1]:ip[r1 + 1]] + 1e-10 * vals[ip[r2]:ip[r2 + 1]]
    s.set_values(v)
    s.factor(prec)
    near = s.pivot_report(prec)
    print("[near-singular] prec %d regular %s shifted %s" % (prec, reg, near), flush=True)
    assert reg["zero_pivots"] == 0 and reg["tiny_pivots"] == 0
    assert near["zero_pivots"] == 0 and near["tiny_pivots"] >= 1
    assert near["min_pivot_ratio"] < 1e-3 * reg["min_pivot_ratio"]
    if prec == 32:
        A2 = SO.csr(A.indptr, A.indices, v, base=0)
        B, _ = rhs_block(A2, 2, 61)
        X, st = s.solve(B, prec=32, tol=1e-8, max_iter=20)
        r = SO.rel_residual(A2, X, B)
        assert st["converged"] == 1 and r <= 1e-8
        assert abs(r - st["rel_residual"]) <= 1e-6 * r
        with pytest.raises(NotConverged) as ei:
            s.solve(B, prec=32, tol=1e-17, max_iter=20)
        st = ei.value.stats
        r = SO.rel_residual(A2, ei.value.x, B)
        assert st["converged"] == 0 and st["rel_residual"] > 1e-17
        # near the FP64 floor b - A x cancels: either residual carries an
        # error of a few eps in units of |b| (observed 7e-17), whatever the
        # order in which the terms are summed
        assert abs(r - st["rel_residual"]) <= 1e-6 * r + 1e-15
    s.close()


def test_fp32_factor_as_accurate_as_dense_partial_pivoting():
    """The complex64 multifrontal factor (pivoting inside 64-row tiles, 3M
    products, tile inverses) against a dense complex64 LU with partial
    pivoting over whole columns (LAPACK through scipy) of the same
    row-equilibrated operator: the first solve's residual, measured in the
    equilibrated rows (powers of two, as the solver scales them), is of the
    same size.  An all-PML 10^3 model, where the unscaled row norms span six
    orders of magnitude."""
    import scipy.linalg as sla
    n = 10
    prm, lam, mu, rho, ir, jc, vals = problem(n, n, n, dx=200.0, freq=0.3, seed=3, perturb=0.05, pmld=2000.0)
    A = SO.csr(ir, jc, vals)
    Ad = A.toarray()
    m = np.maximum(np.abs(Ad.real), np.abs(Ad.imag)).max(axis=1)
    _, ex = np.frexp(m)
    D = np.ldexp(1.0, 1 - ex)
    DA = Ad * D[:, None]
    rng = np.random.default_rng(17)
    B = rng.standard_normal((A.shape[0], 4)) + 1j * rng.standard_normal((A.shape[0], 4))
    lu = sla.lu_factor(DA.astype(np.complex64), check_finite=False)
    Xc = sla.lu_solve(lu, (D[:, None] * B).astype(np.complex64), check_finite=False).astype(np.complex128)

    def scaled_res(X):
        R = D[:, None] * (B - Ad @ X)
        return np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(D[:, None] * B, axis=0))

    s = Solver(n, n, n)
    s.set_model(lam, mu, rho)
    s.assemble(W._lib.WtGridParams(n, n, n, prm["ipml"], 1, 1, prm["dx"], prm["pmld"], prm["pmlf"], prm["omega"],
                                   0.0))
    s.factor(32)
    X, st = s.solve(B, prec=32, tol=0.0, max_iter=1)
    s.close()
    r_dense, r_gpu = scaled_res(Xc), scaled_res(X)
    assert r_dense < 1e-4
    assert r_gpu <= 10 * r_dense, (r_gpu, r_dense)


@pytest.mark.parametrize("nrhs", [64, 40, 33, 32, 17, 16, 5])
def test_wide_block_residual_matches_numpy(nrhs):
    """Blocks of more than 16 columns take the node-wise residual kernel,
    16 and fewer the row-wise one: the FP64 residual the refinement reports
    (and stops on) agrees with numpy's on the returned solution, and the
    block solves each column as it would alone."""
    s, A, vals = _setup(6, 5, 7, 64, seed=8)
    B, Xt = rhs_block(A, nrhs, 41)
    s.factor(32)
    X, st = s.solve(B, prec=32, tol=0.0, max_iter=2)
    r_np = SO.rel_residual(A, X, B)
    assert abs(st["rel_residual"] - r_np) <= 1e-6 * r_np + 1e-15
    x3, _ = s.solve(B[:, 3], prec=32, tol=0.0, max_iter=2)
    assert np.array_equal(x3, X[:, 3])
    s.close()


@pytest.mark.parametrize("nrhs", [130, 65])
def test_more_than_64_columns_refined_in_blocks(nrhs):
    """More than 64 right-hand sides are refined in blocks of at most 64, the
    last one ragged: with a fixed number of solves every column comes out as
    it does alone, host and device buffers give the same block, and refined
    to a tolerance every column meets it (the statistics are the worst
    block's)."""
    s, A, vals = _setup(6, 5, 7, 64, seed=9)
    B, Xt = rhs_block(A, nrhs, 43)
    s.factor(32)
    X3, st3 = s.solve(B, prec=32, tol=0.0, max_iter=3)
    assert st3["nrhs"] == nrhs and st3["iterations"] == 3
    for j in (0, 63, 64, nrhs - 1):
        xj, _ = s.solve(B[:, j], prec=32, tol=0.0, max_iter=3)
        assert np.array_equal(xj, X3[:, j]), j
    L = W.lib()
    dB, dX = DeviceBuffer.from_host(np.ascontiguousarray(B)), DeviceBuffer(B.nbytes)
    st = W._lib.WtSolveStats()
    assert L.wt_solver_solve(s._h, 32, nrhs, dB.ptr, dX.ptr, 1, 0.0, 3, ctypes.byref(st)) == 0
    assert np.array_equal(dX.to_host(np.complex128, B.shape), X3)
    X, st = s.solve(B, prec=32, tol=1e-10, max_iter=30)
    assert st["converged"] == 1 and st["rel_residual"] <= 1e-10
    assert np.max(SO.rel_residual(A, X, B)) <= 1e-10
    s.close()


def test_refinement_reports_non_convergence():
    """Refinement that cannot reach its tolerance ends with WT_ERR_NOCONV
    (NotConverged here) after delivering the solution it reached and its
    statistics; tol <= 0 runs a fixed number of solves without it."""
    from wtomo3d_amd._lib import NotConverged
    s, A, vals = _setup(10, 9, 11, 64, seed=5)
    B, Xt = rhs_block(A, 3, 51)
    s.factor(32)
    # below the FP64 floor: the residual stops falling before max_iter
    with pytest.raises(NotConverged) as ei:
        s.solve(B, prec=32, tol=1e-18, max_iter=40)
    st = ei.value.stats
    assert st["converged"] == 0 and (st["stalled"] == 1 or st["iterations"] == 40)
    assert st["rel_residual"] <= 1e-12
    assert SO.rel_residual(A, ei.value.x, B) <= 1e-12
    # max_iter reached above tol
    with pytest.raises(NotConverged) as ei:
        s.solve(B, prec=32, tol=1e-12, max_iter=1)
    assert ei.value.stats["iterations"] == 1 and ei.value.stats["stalled"] == 0
    X, st = s.solve(B, prec=32, tol=0.0, max_iter=2)
    assert st["converged"] == 1 and st["iterations"] == 2
    X, st = s.solve(B, prec=32, tol=1e-10, max_iter=30)
    assert st["converged"] == 1 and st["stalled"] == 0
    s.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_traction_free_top_solves(prec):
    """The traction-free top (usemin = false, TPFREE) assembled on the GPU,
    factored and solved: against SuperLU of the oracle's operator."""
    import oracle_lib as O
    nx, ny, nz, dx = 10, 9, 11, 250.0
    vp, vs, rho = __import__("solve_cases").layered_model(nx, ny, nz, dx, 6, 0.05)
    lam, mu = O.lame(vp, vs, rho, 64)
    ipml, pmlf = O.pml_params(dx, 2 * dx, 1e-3, 64)
    om = 2 * np.pi * 0.3
    rc, vals = O.assemble(nx, ny, nz, dx, om, lam, mu, rho.ravel(), 2 * dx, pmlf, ipml, usemin=False, prec=64)
    assert rc == 0
    A = SO.csr(*O.graph(nx, ny, nz), vals)
    s = Solver(nx, ny, nz)
    s.set_model(lam, mu, rho.ravel())
    s.assemble(W._lib.WtGridParams(nx, ny, nz, ipml, 1, 0, dx, 2 * dx, pmlf, om, 0.0))
    assert np.abs(s.values() - vals).max() <= 1e-13 * np.abs(vals).max()
    B, Xt = rhs_block(A, 3, 61)
    Xo = SO.solve_superlu(A, B)
    s.factor(prec)
    X, st = s.solve(B, prec=prec, tol=1e-10, max_iter=30)
    s.close()
    assert SO.rel_residual(A, X, B) <= 1e-10
    err_oracle = np.max(np.abs(Xo - Xt)) / np.max(np.abs(Xt))
    assert np.max(np.abs(X - Xt)) / np.max(np.abs(Xt)) <= 10 * err_oracle + 1e-11


@pytest.mark.parametrize("prec", [32, 64])
def test_split_k_back_substitution(prec):
    """The back substitution's update products split along K (here forced
    on every level with pieces >= 16 deep) solve like the whole products:
    to rounding in one solve, and each column the same alone and inside a
    block, since the split depends on the tree and not on the block."""
    s, A, vals = _setup(12, 11, 10, 64, seed=12)
    B, Xt = rhs_block(A, 6, 71)
    s.factor(prec)
    s.set_splitk(0)
    X0, _ = s.solve(B, prec=prec, tol=0.0, max_iter=1)
    s.set_splitk(16)
    X1, st = s.solve(B, prec=prec, tol=0.0, max_iter=1)
    # the split sums the products in another order: the two solutions differ
    # by rounding, amplified by the operator's conditioning exactly as their
    # errors against the true solution are, so compare on that scale
    d = np.max(np.abs(X1 - X0)) / np.max(np.abs(Xt))
    e0 = np.max(np.abs(X0 - Xt)) / np.max(np.abs(Xt))
    e1 = np.max(np.abs(X1 - Xt)) / np.max(np.abs(Xt))
    assert e1 <= 2 * e0 + 1e-12 and d <= 2 * e0 + 1e-12, (d, e0, e1)
    if prec == 64:
        assert d <= 1e-12, d
    assert SO.rel_residual(A, X1, B) <= (1e-2 if prec == 32 else 1e-11)
    x2, _ = s.solve(B[:, 2], prec=prec, tol=0.0, max_iter=1)
    assert np.array_equal(x2, X1[:, 2])
    Xr, st = s.solve(B, prec=prec, tol=1e-10, max_iter=30)
    assert st["rel_residual"] <= 1e-10 and SO.rel_residual(A, Xr, B) <= 1e-10
    s.close()
