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
    e3 = serving.Engine(hbm_budget=50e9, slots=3, tol=1e-10, return_solutions=True, group_max=1)
    e3.run(mix, realtime=False)
    e3.close()
    for a, b in zip(alone, mix):
        assert a.error is None and b.error is None, (a.error, b.error)
        assert a.stats["rel_residual"] <= 1e-10 and b.stats["rel_residual"] <= 1e-10
        # each request runs on its own handle with deterministic kernels
        assert np.array_equal(a.x, b.x)
    assert alone[3].reused  # same model and frequency as request 0, run after it


def test_small_models_batched_equal_alone():
    """Pending small models share one batch (one level-by-level
    factorization and one refined block solve); each result agrees with the
    same request solved alone on its own handle, and a later request reuses
    the batch's factor."""
    reqs = _requests()
    alone = serving.Engine(hbm_budget=50e9, slots=1, tol=1e-10, return_solutions=True, lookahead=1)
    alone.run(reqs, realtime=False)
    alone.close()
    grp = _requests()
    e = serving.Engine(hbm_budget=50e9, slots=2, tol=1e-10, return_solutions=True, group_max=8, group_nodes=48 ** 3)
    e.run(grp, realtime=False)
    assert e.groups >= 1 and e.accounted()
    assert sum(r.batched for r in grp) >= 2
    e.close()
    assert e.used == 0
    for a, b in zip(reqs, grp):
        assert a.error is None and b.error is None, (a.error, b.error)
        assert b.stats["converged"] == 1 and b.stats["rel_residual"] <= 1e-10
        d = np.linalg.norm(a.x - b.x) / np.linalg.norm(a.x)
        assert d <= 1e-7, (b.rid, b.batched, d)
        if not b.batched:
            assert np.array_equal(a.x, b.x)
    assert any(r.reused for r in grp)


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


def _write_project(path, nx, ny, nz, dx, freqs, vp, vs, rho, extra=""):
    from wtomo3d_amd import io
    with open(path + ".ini", "w") as f:
        f.write("[model]\nnx = %d\nny = %d\nnz = %d\ndx = %g\nnom = %d\n" % (nx, ny, nz, dx, len(freqs)))
        for i, fr in enumerate(freqs):
            f.write("freq_%d = %g\n" % (i + 1, fr))
        f.write("[pml]\npmld = %g\n[sources]\nns = 1\n%s" % (3 * dx, extra))
    io.write_model_h5(path + "_model.h5", vp, vs, rho)


def test_on_disk_project_requests(tmp_path):
    """Requests that name an on-disk project (ini + HDF5 model, the
    reference's READINI_FORWARD / MODELIO_READ_MODEL inputs): the model, grid,
    PML and frequencies come from the files, the background is the model's
    lateral average.  The total field satisfies A u = A_bg u_bg (the
    scattered-field identity); a layered project has no scattered field."""
    from wtomo3d_amd import io, media, rhs
    from wtomo3d_amd.solver import Solver
    nx, ny, nz, dx = 14, 12, 16, 250.0
    vp1, vs1, rho1 = media.layered(nx, ny, nz, dx, seed=21)
    d = media.gaussian_field((ny, nx, nz), dx, 1000.0, 0.05, 3)
    het = tmp_path / "het"
    _write_project(str(het), nx, ny, nz, dx, [0.2, 0.35], vp1 * (1 + d), vs1 * (1 + d), rho1 * (1 + 0.3 * d))
    lay = tmp_path / "lay"
    full = [np.broadcast_to(a[None, None, :], (ny, nx, nz)) for a in (vp1, vs1, rho1)]
    _write_project(str(lay), nx, ny, nz, dx, [0.3], *full)
    reqs = serving.project_requests(str(het), 5, 100) + serving.project_requests(str(lay), 3, 200, start_id=2)
    e = serving.Engine(hbm_budget=50e9, slots=2, tol=1e-9, return_solutions=True)
    e.run(reqs, realtime=False)
    for r in reqs:
        assert r.error is None, r.error
        assert r.stats["converged"] == 1 and r.stats["rel_residual"] <= 1e-9
        prob = e._problem(r)
        ain, phi, is_s = e.sources(r)
        om = 2 * np.pi * r.freq
        pr = rhs.profiles(prob.layers, ain, is_s, om, nz, dx, 0.0, 64)
        n = 3 * nx * ny * nz
        ub = rhs.background_fields(nx, ny, nz, dx, pr, phi).to_host(np.complex128, (n, r.nrhs))
        with Solver(nx, ny, nz) as s, Solver(nx, ny, nz) as sb:
            s.set_model(*prob.model)
            s.assemble(prob.params(r.freq))
            sb.set_model(*prob.bg)
            sb.assemble(prob.params(r.freq))
            lhs, rhs_ = s.apply(r.x), sb.apply(ub)
            src = rhs_ - s.apply(ub)
        if r.project == str(lay):
            assert np.abs(src).max() == 0.0  # model == background: no source
            assert np.abs(r.x - ub).max() <= 1e-12 * np.abs(ub).max()
        else:
            assert np.linalg.norm(src) > 1e-3 * np.linalg.norm(rhs_)
            dev = np.linalg.norm(lhs - rhs_, axis=0) / np.linalg.norm(src, axis=0)
            assert dev.max() <= 1e-8, dev
    # the project's parameters reached the assembly
    p = io.read_ini(str(het) + ".ini")
    assert e._problem(reqs[0]).pmld == p["pmld"] == 750.0
    e.close()
    assert e.used == 0
