"""CPU direct-solve oracle -- TEST INFRASTRUCTURE ONLY.

The reference factors and solves with STRUMPACK (strumpack_finter.c:84,102),
a third-party multifrontal LU that is neither vendored nor installed here, so
no reference solution exists to compare against ("parity unpinned" against
the reference solve).  This oracle is an independent FP64 direct solve of
the same CSR system: SuperLU (scipy.sparse.linalg.splu, complex128) on small
grids, and dense LAPACK (numpy.linalg.solve) on tiny ones.  It is pinned by
its own residual and by the agreement of the two solvers in the tests.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def csr(rowptr, colidx, vals, base=1):
    n = rowptr.size - 1
    return sp.csr_matrix((vals, colidx - base, rowptr - base), shape=(n, n))


def solve_superlu(A, B):
    lu = spla.splu(sp.csc_matrix(A, dtype=np.complex128), permc_spec="MMD_AT_PLUS_A")
    return lu.solve(np.asarray(B, np.complex128))


def solve_dense(A, B):
    return np.linalg.solve(A.toarray().astype(np.complex128), np.asarray(B, np.complex128))


def rel_residual(A, X, B):
    """max over columns of ||B - A X|| / ||B||."""
    R = np.asarray(B) - A @ X
    R = R.reshape(R.shape[0], -1)
    B = np.asarray(B).reshape(R.shape)
    return float(np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)))
