"""Serving loop: results are the same alone and in a concurrent mix,
factors are reused, admission respects the HBM budget."""
import numpy as np
import pytest

from wtomo3d_amd import serving

pytestmark = pytest.mark.gpu


def _requests():
    reqs = []
    dims = [(12, 13, 14), (16, 12, 12), (14, 14, 10), (12, 13, 14), (18, 12, 11), (13, 17, 12)]
    for i, d in enumerate(dims):
        key = 7 if i in (0, 3) else 100 + i
        f = 0.3 if i in (0, 3) else 0.1 + 0.05 * i
        reqs.append(serving.Request(i, key, d, 250.0, f, 1 + (3 * i) % 5, 1000 + i, 0.0))
    return reqs


def test_alone_equals_mix_and_reuse():
    alone = _requests()
    e1 = serving.Engine(hbm_budget=50e9, slots=1, tol=1e-10, return_solutions=True, lookahead=1)
    e1.run(alone, realtime=False)
    e1.close()
    mix = _requests()
    e3 = serving.Engine(hbm_budget=50e9, slots=3, tol=1e-10, return_solutions=True)
    e3.run(mix, realtime=False)
    e3.close()
    for a, b in zip(alone, mix):
        assert a.error is None and b.error is None, (a.error, b.error)
        assert a.stats["rel_residual"] <= 1e-10 and b.stats["rel_residual"] <= 1e-10
        # each request runs on its own handle with deterministic kernels
        assert np.array_equal(a.x, b.x)
    assert alone[3].reused  # same model and frequency as request 0, run after it


def test_admission_budget():
    reqs = _requests()[:3]
    need = max(serving.Engine()._estimate(r.dims) for r in reqs)
    e = serving.Engine(hbm_budget=need * 1.05, slots=3, tol=1e-8, keep_factors=False)
    wall = e.run(reqs, realtime=False)
    e.close()
    rep = serving.latency_report(reqs, wall)
    assert rep["completed"] == 3 and rep["errors"] == 0
    tiny = serving.Engine(hbm_budget=1e6, slots=2)
    r = _requests()[:1]
    tiny.run(r, realtime=False)
    assert r[0].error is not None  # cannot fit: rejected, not deadlocked
