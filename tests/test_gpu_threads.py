"""Library calls from several host threads while another thread captures
graphs: a batch's refinement step is captured into a graph on every change
of its column count, and meanwhile other threads create and free solver
handles, copy synchronously and launch on the null stream -- the calls HIP
refuses while a capture is open on any stream.  No call may fail, and the
captured solves must equal the same solves made with no other thread
running."""
import threading

import numpy as np
import pytest

import wtomo3d_amd as W
from wtomo3d_amd import batch, media, rhs
from wtomo3d_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _batch():
    ms = batch.small_models(5, seed=77, side=(12, 18), nrhs=(1, 4))
    b = batch.Batch([m["dims"] for m in ms])
    b.set_graphs(True)
    lams, mus, rhos = zip(*(media.lame(m["vp"], m["vs"], m["rho"]) for m in ms))
    b.set_models(list(lams), list(mus), list(rhos))
    b.assemble([m["params"] for m in ms])
    b.factor(32)
    bs = b.apply([batch.plane_wave_block(m) for m in ms])
    return b, [x.reshape(x.shape[0], -1)[:, :1] for x in bs]


def _widths():
    return [1 + it % 4 for it in range(10)]


def _solves(b, b1):
    out = []
    for c in _widths():
        xs, st = b.solve([np.repeat(x, c, axis=1) for x in b1], prec=32, tol=1e-9, max_iter=20)
        assert all(s["converged"] == 1 for s in st)
        out.append(xs)
    return out


def _churn(stop, errors, seed):
    rng = np.random.default_rng(seed)
    try:
        while not stop.is_set():
            n = int(rng.integers(10, 16))
            vp = np.full((n, n, n), 5000.0)
            lam, mu, rho = media.lame(vp, vp / 1.8, vp * 0 + 2500.0)
            with Solver(n, n, n) as s:
                s.set_model(lam, mu, rho)
                s.assemble(W.grid_params(n, n, n, 200.0, 2 * np.pi * 0.3, pmld=1000.0))
                v = s.values()
                assert np.all(np.isfinite(v))
            a = np.arange(4096, dtype=np.complex128)
            d = rhs.DeviceBuffer.from_host(a)
            assert np.array_equal(d.to_host(np.complex128, a.shape), a)
            d.free()
            W.graph(n, n, n)
    except Exception as ex:  # reported by the main thread
        errors.append(repr(ex))


@pytest.mark.timeout(300)
def test_graph_capture_beside_synchronous_calls():
    b, b1 = _batch()
    try:
        print("[threads] serial solves", flush=True)
        ref = _solves(b, b1)
        stop, errors = threading.Event(), []
        workers = [threading.Thread(target=_churn, args=(stop, errors, k)) for k in range(2)]
        for t in workers:
            t.start()
        try:
            print("[threads] the same solves beside two churning threads", flush=True)
            got = _solves(b, b1)
        finally:
            stop.set()
            for t in workers:
                t.join()
    finally:
        b.close()
    assert not errors, errors
    for xa, xb in zip(ref, got):
        for u, v in zip(xa, xb):
            assert np.array_equal(u, v)
