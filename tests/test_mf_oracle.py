"""The CPU multifrontal restatement: dense partial LU against a full LU,
and its nested-dissection statistics against closed forms."""
import os
import sys

import numpy as np
import scipy.linalg as sla

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mf_oracle as MF  # noqa: E402


def test_partial_lu_schur_complement():
    rng = np.random.default_rng(0)
    p, u = 40, 25
    m = p + u
    F = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    lu, piv, CB = MF.front_partial_lu(F, p)
    S = F[p:, p:] - F[p:, :p] @ np.linalg.solve(F[:p, :p], F[:p, p:])
    assert np.allclose(CB, S, rtol=1e-10, atol=1e-10)


def test_nd_tree_partitions_the_grid():
    for dims, leaf in (((24, 24, 24), 64), ((9, 22, 7), 27), ((96, 96, 64), 64)):
        fr = MF.nd_fronts(*dims, leaf=leaf)
        assert sum(p for p, _, _ in fr) == 3 * dims[0] * dims[1] * dims[2]
        root = fr[-1]
        assert root[1] == 0 and root[2] == 0
        assert all(u % 3 == 0 and p % 3 == 0 for p, u, _ in fr)


def test_nd_root_separator_c3():
    fr = MF.nd_fronts(96, 96, 64)
    assert fr[-1][0] == 3 * 96 * 64  # plane across y (ties prefer y)


def test_library_symbolic_matches_restatement():
    """The library's host symbolic analysis builds the same tree (front
    count, depth, largest front, total factorization flops)."""
    import ctypes
    from wtomo3d_amd import lib
    for dims, leaf in (((24, 24, 24), 64), ((9, 22, 7), 27), ((96, 96, 64), 64), ((17, 40, 23), 100)):
        info = (ctypes.c_int64 * 8)()
        fl = ctypes.c_double()
        assert lib().wt_symbolic_info(*dims, leaf, info, ctypes.byref(fl)) == 0
        fr = MF.nd_fronts(*dims, leaf=leaf)
        assert info[2] == len(fr)
        assert info[3] == max(l for _, _, l in fr) + 1
        assert info[4] == max(p for p, _, _ in fr)
        assert info[5] == max(u for _, u, _ in fr)
        assert abs(fl.value - MF.tree_flops(fr)) <= 1e-9 * fl.value
        assert info[6] == sum(p * p + 2 * p * u for p, u, _ in fr)
