"""Serving loop bookkeeping on the CPU: the factor cache, the HBM
accounting and the error paths, with the GPU work replaced by stand-ins
(the engine's _factor / _solve / _slot hooks)."""
import threading
import time

import numpy as np
import pytest

from wtomo3d_amd import media, serving
from wtomo3d_amd._lib import NotConverged


class FakeSolver:
    def __init__(self, log):
        self.closed = False
        self.log = log

    def close(self):
        assert not self.closed, "handle closed twice"
        self.closed = True
        self.log.append("close")


class MockEngine(serving.Engine):
    """Engine whose factorizations and solves are timed sleeps; requests in
    ``fail`` raise in their solve, those in ``noconv`` end unconverged."""

    def __init__(self, fail=(), noconv=(), factor_s=0.05, solve_s=0.01, nbytes=1000, **kw):
        kw.setdefault("group_max", 1)
        super().__init__(**kw)
        self.fail, self.noconv = set(fail), set(noconv)
        self.factor_s, self.solve_s, self.nbytes = factor_s, solve_s, nbytes
        self.log = []
        self.lock = threading.Lock()
        self.max_used = 0

    def _estimate(self, dims):
        return self.nbytes

    def _problem(self, r):
        return None

    def _factor(self, r, prob):
        time.sleep(self.factor_s)
        with self.lock:
            self.log.append(("factor", r.rid))
        return FakeSolver(self.log)

    def _solve(self, r, s, prob, slot):
        if s.closed:
            raise RuntimeError("closed handle")
        with self.cv:
            self.max_used = max(self.max_used, self.used)
            assert self.used == sum(e.nbytes for e in self.cache.values())
        time.sleep(self.solve_s)
        if r.rid in self.fail:
            raise RuntimeError("injected failure")
        r.stats = {"rel_residual": 0.0, "iterations": 1}
        if r.rid in self.noconv:
            r.stats["rel_residual"] = 1e-3
            raise NotConverged("request %d" % r.rid, r.stats)

    def _slot(self):
        return (lambda *a: None), (lambda: None)

    def _factor_group(self, rs, probs):
        time.sleep(self.factor_s)
        with self.lock:
            self.log.append(("group", tuple(r.rid for r in rs)))
        return FakeSolver(self.log)

    def _solve_group(self, rs, e, slot):
        if e.batch.closed:
            raise RuntimeError("closed batch")
        with self.cv:
            self.max_used = max(self.max_used, self.used)
        time.sleep(self.solve_s)
        for r in rs:
            assert r.key() in e.index
            r.stats = {"rel_residual": 0.0, "iterations": 1}
            if r.rid in self.fail:
                r.error = "injected failure"


def req(rid, key, freq=0.3, dims=(32, 32, 32)):
    return serving.Request(rid, key, dims, 200.0, freq, 4, rid, 0.0)


def factors(e):
    return [x for x in e.log if isinstance(x, tuple) and x[0] == "factor"]


def test_overlapping_same_key_factor_once():
    e = MockEngine(slots=2, hbm_budget=10_000)
    rs = [req(0, 7), req(1, 7)]
    e.run(rs, realtime=False)
    assert all(r.error is None for r in rs)
    assert len(factors(e)) == 1 and e.factorizations == 1
    assert sum(r.reused for r in rs) == 1
    assert len(e.cache) == 1 and e.used == 1000 and e.accounted()
    e.close()
    assert e.used == 0 and not e.cache and e.log.count("close") == 1


def test_progress_lines_go_to_stderr(capfd):
    """The engine's progress report stays off stdout, which carries the
    benchmark's single result line."""
    e = MockEngine(slots=2, hbm_budget=10_000, report_every=1, factor_s=0.0, solve_s=0.0)
    e.run([req(i, i % 2) for i in range(4)], realtime=False)
    e.close()
    out, err = capfd.readouterr()
    assert out == "" and err.count("serving: ") == 4


def test_many_overlapping_keys():
    e = MockEngine(slots=4, hbm_budget=100_000, factor_s=0.02, solve_s=0.005)
    rs = [req(i, i % 3, freq=0.1 * (1 + i % 3)) for i in range(12)]
    e.run(rs, realtime=False)
    assert all(r.error is None for r in rs)
    assert e.factorizations == 3 and sum(r.reused for r in rs) == 9
    assert e.used == 3000 and e.accounted()
    e.close()
    assert e.used == 0


def test_failure_on_reused_entry_refactors():
    e = MockEngine(slots=1, fail={1}, hbm_budget=10_000)
    rs = [req(0, 7), req(1, 7), req(2, 7), req(3, 7)]
    e.run(rs, realtime=False)
    assert rs[0].error is None and rs[1].error == "injected failure"
    assert rs[2].error is None and rs[3].error is None
    assert not rs[2].reused and rs[3].reused  # the failed entry was dropped, 2 refactored
    assert e.factorizations == 2
    assert e.used == 1000 and e.accounted()
    e.close()
    assert e.used == 0 and e.log.count("close") == 2


def test_not_converged_is_an_error():
    e = MockEngine(slots=2, noconv={0}, hbm_budget=10_000)
    rs = [req(0, 7), req(1, 8)]
    e.run(rs, realtime=False)
    assert rs[0].error is not None and "request 0" in rs[0].error
    assert rs[1].error is None
    assert len(e.cache) == 1 and e.used == 1000 and e.accounted()
    rep = serving.latency_report(rs, 1.0)
    assert rep["completed"] == 1 and rep["errors"] == 1
    e.close()


def test_eviction_keeps_budget():
    e = MockEngine(slots=2, hbm_budget=2500)
    rs = [req(i, 100 + i) for i in range(6)]
    e.run(rs, realtime=False)
    assert all(r.error is None for r in rs)
    assert e.max_used <= 2500 and e.used <= 2500 and e.accounted()
    assert len(e.cache) <= 2
    e.close()
    assert e.used == 0


def test_no_keep_releases_everything():
    e = MockEngine(slots=3, hbm_budget=10_000, keep_factors=False)
    rs = [req(i, i % 2) for i in range(6)]
    e.run(rs, realtime=False)
    assert all(r.error is None for r in rs)
    assert e.used == 0 and not e.cache
    assert e.factorizations == 6


def test_oversized_request_rejected():
    e = MockEngine(slots=2, hbm_budget=500)
    rs = [req(0, 1), req(1, 2)]
    e.run(rs, realtime=False)
    assert all(r.error is not None and "budget" in r.error for r in rs)
    assert e.used == 0


def test_background_1d_layered_recovers_layers():
    nx, ny, nz, dx = 6, 5, 40, 200.0
    vp1, vs1, rho1 = media.layered(nx, ny, nz, dx, seed=3)
    full = [np.broadcast_to(a[None, None, :], (ny, nx, nz)) for a in (vp1, vs1, rho1)]
    a, b, c, lay = media.background_1d(*full, dx)
    assert np.array_equal(a, vp1) and np.array_equal(b, vs1) and np.array_equal(c, rho1)
    # every node's layer by the tops is the layer that holds its values
    z = np.arange(nz) * dx
    k = np.searchsorted(lay["top"], z, side="right") - 1
    assert np.array_equal(lay["vp"][k], vp1) and np.array_equal(lay["rho"][k], rho1)
    assert lay["top"][0] == 0.0 and np.all(np.diff(lay["top"]) > 0)
    assert not np.any(np.isin(lay["top"][1:], z))  # no node on an interface


def test_background_1d_limits_and_homogeneous():
    rng = np.random.default_rng(1)
    ny, nx, nz, dx = 4, 3, 100, 100.0
    vp = 5000 + rng.random((ny, nx, nz))
    a, b, c, lay = media.background_1d(vp, vp / 1.7, 2500 + 0 * vp, dx)
    assert lay["vp"].size <= 64 and a.size == nz
    z = np.arange(nz) * dx
    k = np.searchsorted(lay["top"], z, side="right") - 1
    assert np.array_equal(lay["vp"][k], a)
    # the per-depth mean is kept within each run
    assert abs(a.mean() - vp.mean()) < 1e-9 * vp.mean() + 1e-6
    h = np.full((ny, nx, 10), 6000.0)
    a, b, c, lay = media.background_1d(h, h / np.sqrt(3), h * 0 + 2700.0, 500.0)
    assert lay["vp"].size == 2 and lay["top"][1] >= 10 * 500.0
    assert np.all(a == 6000.0)


def test_project_requests(tmp_path):
    pytest.importorskip("ctypes")
    from wtomo3d_amd import io
    name = str(tmp_path / "proj")
    with open(name + ".ini", "w") as f:
        f.write("[model]\nnx = 12\nny = 11\nnz = 10\ndx = 250\nnom = 2\nfreq_1 = 0.2\nfreq_2 = 0.3\n"
                "[sources]\nns = 1\n")
    try:
        vp = np.full((11, 12, 10), 5000.0)
        io.write_model_h5(name + "_model.h5", vp, vp / 1.73, vp * 0 + 2600.0)
    except io.ProjectError as ex:
        pytest.skip(str(ex))
    rs = serving.project_requests(name, 3, 10)
    assert [r.freq for r in rs] == [0.2, 0.3] and all(r.project == name for r in rs)
    prob = serving.Engine._load_project(name)
    assert prob.dims == (12, 11, 10) and prob.dx == 250.0 and prob.usemin and prob.fst
    assert rs[0].key() != rs[1].key()


def test_small_models_share_one_batch():
    """Pending new small models are admitted together into one batch entry
    that serves every member's key; later requests for those keys reuse it,
    and its bytes are held once."""
    e = MockEngine(slots=1, hbm_budget=100_000, group_max=4, group_nodes=48 ** 3, lookahead=8)
    rs = [req(i, 10 + i) for i in range(4)] + [req(4, 11), req(5, 13)]
    e.run(rs, realtime=False)
    assert all(r.error is None for r in rs)
    groups = [x for x in e.log if isinstance(x, tuple) and x[0] == "group"]
    assert groups == [("group", (0, 1, 2, 3))] and e.groups == 1 and e.factorizations == 4
    assert rs[4].reused and rs[5].reused
    assert len(e.cache) == 4 and len({id(v) for v in e.cache.values()}) == 1
    assert e.used == 4000 and e.accounted()
    e.close()
    assert e.used == 0 and not e.cache and e.log.count("close") == 1


def test_large_models_are_not_grouped():
    e = MockEngine(slots=2, hbm_budget=100_000, group_max=4, group_nodes=48 ** 3)
    rs = [req(i, 20 + i, dims=(64, 48, 48)) for i in range(3)]
    e.run(rs, realtime=False)
    assert not [x for x in e.log if isinstance(x, tuple) and x[0] == "group"]
    assert e.factorizations == 3 and e.accounted()
    e.close()


def test_failed_group_member_drops_the_batch_once():
    e = MockEngine(slots=1, hbm_budget=100_000, group_max=3, fail={1})
    rs = [req(i, 30 + i) for i in range(3)] + [req(3, 30)]
    e.run(rs, realtime=False)
    assert rs[1].error == "injected failure" and rs[0].error is None and rs[2].error is None
    assert not rs[3].reused  # the batch was dropped with its failed member: key 30 refactored
    assert e.accounted() and e.log.count("close") == 1
    e.close()
    assert e.used == 0


def test_group_eviction_under_budget():
    e = MockEngine(slots=1, hbm_budget=3500, group_max=3)
    rs = [req(i, 40 + i) for i in range(6)]
    e.run(rs, realtime=False)
    assert all(r.error is None for r in rs)
    assert e.max_used <= 3500 and e.accounted()
    e.close()
    assert e.used == 0
