"""The CPU solve oracle agrees with itself (SuperLU vs dense LAPACK) and
has FP64-level residuals on the elastic operator."""
import sys
import os

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import solve_oracle as SO  # noqa: E402
from solve_cases import problem, rhs_block  # noqa: E402


def test_superlu_matches_dense_lapack():
    prm, lam, mu, rho, ir, jc, vals = problem(5, 4, 6, seed=3)
    A = SO.csr(ir, jc, vals)
    B, Xt = rhs_block(A, 3, 4)
    X1 = SO.solve_superlu(A, B)
    X2 = SO.solve_dense(A, B)
    assert np.max(np.abs(X1 - X2)) <= 1e-10 * np.max(np.abs(X2))
    assert SO.rel_residual(A, X1, B) <= 1e-12


def test_superlu_residual_small_grid():
    prm, lam, mu, rho, ir, jc, vals = problem(10, 9, 11, seed=5)
    A = SO.csr(ir, jc, vals)
    B, Xt = rhs_block(A, 2, 6)
    X = SO.solve_superlu(A, B)
    assert SO.rel_residual(A, X, B) <= 1e-12
    assert np.max(np.abs(X - Xt)) <= 1e-8 * np.max(np.abs(Xt))
