"""CPU multifrontal restatement for the baseline -- TEST INFRASTRUCTURE ONLY.

STRUMPACK's factorization (strumpack_finter.c:84) is not available here.
This module restates the algorithm the path runs -- nested dissection of the
grid into an elimination tree of dense fronts, each partially factored (LU
with partial pivoting of the pivot block, L21 / U12 panel solves and the
Schur complement CB = F22 - L21 U12) -- with numpy/LAPACK on the host:

- ``nd_fronts``: the geometric nested dissection (longest axis cut by a
  one-node plane at its middle, leaves of at most `leaf` nodes, update set
  = one-node halo of the box), giving every front's (pivots, update size).
- ``front_partial_lu``: the dense partial factorization of one front.
- ``cpu_rate_sample``: times that factorization on fronts of the tree
  (largest first) within a time budget and returns the measured flop rate,
  from which the whole-tree CPU time is extrapolated.
Parity of the tree statistics with the GPU library's symbolic analysis is
tested (tests/test_mf_oracle.py); the sampled arithmetic is LAPACK's.
"""
import time

import numpy as np
import scipy.linalg as sla


def nd_fronts(nx, ny, nz, leaf=64):
    """[(p, u, level)] in post-order."""
    out = []

    def rec(b, depth):
        x0, x1, y0, y1, z0, z1 = b
        dims = [x1 - x0, y1 - y0, z1 - z0]
        vol = dims[0] * dims[1] * dims[2]
        ax = 1
        if dims[0] > dims[ax]:
            ax = 0
        if dims[2] > dims[ax]:
            ax = 2
        if vol > leaf and dims[ax] >= 3:
            lo = b[2 * ax]
            s = lo + dims[ax] // 2
            c0, c1 = list(b), list(b)
            c0[2 * ax + 1] = s
            c1[2 * ax] = s + 1
            rec(tuple(c0), depth + 1)
            rec(tuple(c1), depth + 1)
            piv = vol // dims[ax]
        else:
            piv = vol
        hx = min(x1 + 1, nx) - max(x0 - 1, 0)
        hy = min(y1 + 1, ny) - max(y0 - 1, 0)
        hz = min(z1 + 1, nz) - max(z0 - 1, 0)
        out.append((3 * piv, 3 * (hx * hy * hz - vol), depth))

    rec((0, nx, 0, ny, 0, nz), 0)
    return out


def front_flops(p, u):
    return 8.0 * (p ** 3 / 3.0 + p * p * u + p * u * u)


def tree_flops(fronts):
    return float(sum(front_flops(p, u) for p, u, _ in fronts))


def front_partial_lu(F, p):
    """Partial LU of the dense front F ((p+u) x (p+u)) over its first p
    pivots; returns (lu, piv, CB)."""
    lu, piv = sla.lu_factor(F[:p, :p], check_finite=False)
    F12 = F[:p, p:].copy()
    for i, j in enumerate(piv):  # row interchanges of the pivot block
        if i != j:
            F12[[i, j]] = F12[[j, i]]
    U12 = sla.solve_triangular(lu, F12, lower=True, unit_diagonal=True, check_finite=False)
    L21 = sla.solve_triangular(lu, F[p:, :p].T, trans="T", lower=False, check_finite=False).T
    CB = F[p:, p:] - L21 @ U12
    return lu, piv, CB


def cpu_rate_sample(fronts, dtype=np.complex64, budget_s=20.0, seed=0):
    """Factor fronts of the tree, largest first, until the budget is spent.
    Returns (flops done, seconds, list of (p, u) sampled)."""
    rng = np.random.default_rng(seed)
    order = sorted(set((p, u) for p, u, _ in fronts), key=lambda t: -front_flops(*t))
    done_f, done_t, sample = 0.0, 0.0, []
    # start with the largest front whose predicted time fits the budget
    rate_guess = 2e11
    for p, u in order:
        if front_flops(p, u) / rate_guess > budget_s - done_t:
            continue
        m = p + u
        F = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))).astype(dtype)
        F[np.arange(m), np.arange(m)] += 4.0 * np.sqrt(m)
        t = time.perf_counter()
        front_partial_lu(F, p)
        dt = time.perf_counter() - t
        done_f += front_flops(p, u)
        done_t += dt
        sample.append((p, u))
        rate_guess = done_f / done_t
        if done_t >= 0.5 * budget_s:
            break
    return done_f, done_t, sample


# ---------------------------------------------------------------------------
# FP64 multifrontal direct solve (the reference solution of full-size solve tests)

def nd_tree(nx, ny, nz, leaf=64):
    """The tree of ``nd_fronts`` with node sets: list of fronts in post-order,
    each a dict with ``piv`` (the grid nodes it eliminates: the separator
    plane, or the whole leaf box), ``upd`` (the one-node halo of its box
    outside it) and ``kids`` (indices of its children).  Node g = iz + nz ix
    + nz nx iy (0-based), the reference's unknown order (srgraphs.f:659-665)."""
    out = []

    def nodes(x0, x1, y0, y1, z0, z1):
        iy, ix, iz = np.meshgrid(np.arange(y0, y1), np.arange(x0, x1), np.arange(z0, z1), indexing="ij")
        return (iz + nz * ix + nz * nx * iy).ravel()

    def rec(b):
        x0, x1, y0, y1, z0, z1 = b
        dims = [x1 - x0, y1 - y0, z1 - z0]
        vol = dims[0] * dims[1] * dims[2]
        ax = 1
        if dims[0] > dims[ax]:
            ax = 0
        if dims[2] > dims[ax]:
            ax = 2
        kids = []
        if vol > leaf and dims[ax] >= 3:
            s = b[2 * ax] + dims[ax] // 2
            c0, c1 = list(b), list(b)
            c0[2 * ax + 1] = s
            c1[2 * ax] = s + 1
            kids = [rec(tuple(c0)), rec(tuple(c1))]
            sep = list(b)
            sep[2 * ax], sep[2 * ax + 1] = s, s + 1
            piv = nodes(*sep)
        else:
            piv = nodes(*b)
        halo = nodes(max(x0 - 1, 0), min(x1 + 1, nx), max(y0 - 1, 0), min(y1 + 1, ny), max(z0 - 1, 0),
                     min(z1 + 1, nz))
        upd = np.setdiff1d(halo, nodes(*b))
        out.append(dict(piv=np.sort(piv), upd=upd, kids=kids))
        return len(out) - 1

    rec((0, nx, 0, ny, 0, nz))
    return out


def _unk(nodes):
    return (3 * np.asarray(nodes, np.int64)[:, None] + np.arange(3)).ravel()


class MultifrontalLU:
    """LU of a sparse matrix on the 27-point elastic grid by the multifrontal
    method over ``nd_tree``: each front assembles the original entries of its
    pivot rows and columns and its children's contribution blocks, factors
    its pivot block with LAPACK's partial pivoting (getrf), forms the panels
    L21, U12 and the contribution block F22 - L21 U12.  Complex128; the
    checker of the GPU factorization at sizes where SuperLU is too slow."""

    def __init__(self, A, nx, ny, nz, leaf=64):
        A = A.tocsr()
        self.n = A.shape[0]
        tree = nd_tree(nx, ny, nz, leaf)
        At = A.T.tocsr()  # columns of the pivot unknowns as rows
        self.fronts = []
        cbs = {}
        for f in tree:
            P, U = _unk(f["piv"]), _unk(f["upd"])
            I = np.concatenate([P, U])
            p, m = P.size, I.size
            pos = np.full(self.n, -1, np.int64)
            pos[I] = np.arange(m)
            F = np.zeros((m, m), np.complex128)
            rows = A[P][:, I].toarray()        # pivot rows (columns in the front)
            F[:p, :] += rows
            cols = At[P][:, U].toarray().T     # pivot columns, update rows
            F[p:, :p] += cols
            for k in f["kids"]:
                ku, cb = cbs.pop(k)
                ix = pos[ku]
                F[np.ix_(ix, ix)] += cb
            lu, piv = sla.lu_factor(F[:p, :p], check_finite=False)
            F12 = F[:p, p:]
            for i, j in enumerate(piv):
                if i != j:
                    F12[[i, j]] = F12[[j, i]]
            U12 = sla.solve_triangular(lu, F12, lower=True, unit_diagonal=True, check_finite=False)
            L21 = sla.solve_triangular(lu, F[p:, :p].T, trans="T", lower=False, check_finite=False).T
            if U.size:
                cbs[len(self.fronts)] = (U, F[p:, p:] - L21 @ U12)
            self.fronts.append((P, U, lu, piv, L21, U12))

    def solve(self, B):
        B = np.asarray(B, np.complex128)
        one = B.ndim == 1
        Y = B.reshape(self.n, -1).copy()
        for P, U, lu, piv, L21, U12 in self.fronts:
            y = Y[P]
            for i, j in enumerate(piv):
                if i != j:
                    y[[i, j]] = y[[j, i]]
            y = sla.solve_triangular(lu, y, lower=True, unit_diagonal=True, check_finite=False)
            Y[P] = y
            if U.size:
                Y[U] -= L21 @ y
        for P, U, lu, piv, L21, U12 in reversed(self.fronts):
            y = Y[P]
            if U.size:
                y = y - U12 @ Y[U]
            Y[P] = sla.solve_triangular(lu, y, lower=False, check_finite=False)
        return Y[:, 0] if one else Y
