"""Host side of the batched small-model solver (no GPU): the CSR pattern
generated on the host, and the symbolic analysis of a batch's union.

The host pattern must equal the oracle's SREL3D + SPARSE_ADJ2CRS pattern
(srgraphs.f:638-734, sparse.f:13-49) bit for bit.  The union of independent
models is pinned by additivity: its unknowns, CSR entries, fronts, factor
entries and flop count are the sums of the models' own analyses, its depth
and largest front the maxima."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import wtomo3d_amd as W
from wtomo3d_amd import batch as Bt

P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 4, 5), (7, 5, 6), (12, 11, 10), (1, 3, 2)])
@pytest.mark.parametrize("base", [0, 1])
def test_host_pattern_matches_oracle(dims, base):
    nx, ny, nz = dims
    ir, jc = O.graph(nx, ny, nz)  # 1-based
    r = np.zeros(ir.size, np.int32)
    c = np.zeros(jc.size, np.int32)
    assert W.lib().wt_graph_host(nx, ny, nz, base, P(r), P(c)) == 0
    assert np.array_equal(r, ir - 1 + base) and np.array_equal(c, jc - 1 + base)


def _single(d, leaf=64):
    info = (ctypes.c_int64 * 8)()
    fl = ctypes.c_double()
    assert W.lib().wt_symbolic_info(d[0], d[1], d[2], leaf, info, ctypes.byref(fl)) == 0
    keys = ("n", "nnz", "fronts", "levels", "max_p", "max_u", "factor_len", "cb_len")
    out = {k: int(info[i]) for i, k in enumerate(keys)}
    out["flops"] = fl.value
    return out


def test_union_is_additive():
    dims = [(6, 7, 5), (9, 4, 8), (5, 5, 5), (12, 6, 7), (4, 9, 3)]
    U = Bt.symbolic_info(dims)
    parts = [_single(d) for d in dims]
    for k in ("n", "nnz", "fronts", "factor_len"):
        assert U[k] == sum(p[k] for p in parts), k
    assert U["levels"] == max(p["levels"] for p in parts)
    assert U["max_p"] == max(p["max_p"] for p in parts)
    assert U["flops"] == pytest.approx(sum(p["flops"] for p in parts), rel=1e-12)
    # per-level arenas are shared across models: never larger than the sum
    assert U["cb_len"] <= sum(p["cb_len"] for p in parts)


def test_union_of_one_equals_the_model():
    U = Bt.symbolic_info([(10, 9, 8)])
    S = _single((10, 9, 8))
    for k in ("n", "nnz", "fronts", "levels", "max_p", "factor_len", "cb_len"):
        assert U[k] == S[k], k


def test_small_models_workload_shape():
    ms = Bt.small_models(nmodels=12, seed=1004)
    assert len(ms) == 12
    for m in ms:
        assert all(16 <= d <= 40 for d in m["dims"])
        assert 0.05 <= m["freq"] <= 0.5 and 1 <= m["nrhs"] <= 16
        nx, ny, nz = m["dims"]
        assert m["vp"].shape == (ny, nx, nz)
        vp1 = m["vp"][0, 0, :]
        assert np.all(np.diff(vp1) >= 0) and vp1.min() >= 4000 and vp1.max() <= 8000
        assert m["params"].ipml == 6
    # seeded: the same draw twice
    a, b = Bt.small_models(nmodels=3, seed=7), Bt.small_models(nmodels=3, seed=7)
    assert [m["dims"] for m in a] == [m["dims"] for m in b]


def test_batch_rejects_bad_dims():
    h = ctypes.c_void_p()
    d = np.array([4, 4, 1], np.int32)
    assert W.lib().wt_batch_create(1, P(d), 64, 0, 0, ctypes.byref(h)) != 0
