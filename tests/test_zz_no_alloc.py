"""wt_haskgrn_noalloc (no device allocation, static scratch) against
wt_haskgrn through the C ABI: the same profiles, wavenumbers and statuses
bit for bit, on two calls in a row (the second smaller), in both
precisions, and the size limit refused before any device work."""
import ctypes

import numpy as np
import pytest

from wtomo3d_amd import lib, media
from wtomo3d_amd.rhs import DeviceBuffer

pytestmark = pytest.mark.gpu
_vp = ctypes.c_void_p


def _call(fn, prec, lay, ain, iss, omega, nz, dz, zorig):
    ft = np.float32 if prec == 32 else np.float64
    ct = np.complex64 if prec == 32 else np.complex128
    arrs = [np.ascontiguousarray(lay[k], ft) for k in ("vp", "vs", "rho", "top", "vpd", "vsd")]
    ain = np.ascontiguousarray(ain, ft)
    iss = np.ascontiguousarray(iss, np.int32)
    ns = ain.size
    nb = ns * nz * np.dtype(ct).itemsize
    U, W, K = DeviceBuffer(nb), DeviceBuffer(nb), DeviceBuffer(ns * 8)
    status = np.full(ns, -7, np.int32)
    rc = fn(prec, arrs[0].size, *[a.ctypes.data_as(_vp) for a in arrs], ns, ain.ctypes.data_as(_vp),
            iss.ctypes.data_as(_vp), float(omega), nz, float(dz), float(zorig), U.ptr, W.ptr, K.ptr,
            status.ctypes.data_as(_vp), None)
    out = (rc, U.to_host(ct, (ns, nz)), W.to_host(ct, (ns, nz)), K.to_host(np.float64, (ns,)), status)
    for b in (U, W, K):
        b.free()
    return out


def _same(a, b):
    assert a[0] == b[0] == 0, (a[0], b[0])
    for x, y in zip(a[1:], b[1:]):
        assert x.tobytes() == y.tobytes()


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_equals_haskgrn_twice_then_smaller(prec):
    L = lib()
    rng = np.random.default_rng(3)
    big = media.layer_stack(96, 96, 64, 200.0, 1003)             # 6 layers
    small = media.layer_stack(40, 40, 30, 250.0, 17, nlayers=3)  # 3 layers
    calls = [(big, 200, 64, 200.0, 0.0, 2 * np.pi * 0.3), (small, 37, 30, 250.0, 0.0, 2 * np.pi * 0.15)]
    for lay, ns, nz, dz, zorig, om in calls:
        ain = rng.uniform(np.deg2rad(5.0), np.deg2rad(60.0), ns)
        iss = (rng.random(ns) < 0.5).astype(np.int32)
        ref = _call(L.wt_haskgrn, prec, lay, ain, iss, om, nz, dz, zorig)
        got = _call(L.wt_haskgrn_noalloc, prec, lay, ain, iss, om, nz, dz, zorig)
        _same(got, ref)
        assert np.all(got[4] == 0)


def test_noalloc_status_of_a_depth_above_the_first_layer():
    """The reference's error path (a depth above top[0]: status 1, the
    profile zero from there on) comes back through the staged status."""
    L = lib()
    lay = media.layer_stack(20, 20, 12, 250.0, 5, nlayers=3)
    ain = np.deg2rad([15.0, 25.0, 35.0])
    iss = np.array([0, 1, 0], np.int32)
    ref = _call(L.wt_haskgrn, 64, lay, ain, iss, 2 * np.pi * 0.2, 12, 250.0, -400.0)
    got = _call(L.wt_haskgrn_noalloc, 64, lay, ain, iss, 2 * np.pi * 0.2, 12, 250.0, -400.0)
    assert got[0] == ref[0] == 0
    for x, y in zip(got[1:], ref[1:]):
        assert x.tobytes() == y.tobytes()
    assert np.all(got[4] == 1)
    # and a valid call right after leaves no stale status behind
    ok = _call(L.wt_haskgrn_noalloc, 64, lay, ain, iss, 2 * np.pi * 0.2, 12, 250.0, 0.0)
    assert ok[0] == 0 and np.all(ok[4] == 0)


def test_noalloc_refuses_more_sources_than_its_scratch():
    L = lib()
    lay = media.layer_stack(20, 20, 12, 250.0, 5, nlayers=3)
    ns = 4097
    arrs = [np.ascontiguousarray(lay[k], np.float64) for k in ("vp", "vs", "rho", "top", "vpd", "vsd")]
    ain = np.full(ns, 0.3)
    iss = np.zeros(ns, np.int32)
    status = np.zeros(ns, np.int32)
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = L.wt_haskgrn_noalloc(64, arrs[0].size, *[a.ctypes.data_as(_vp) for a in arrs], ns,
                              ain.ctypes.data_as(_vp), iss.ctypes.data_as(_vp), 1.0, 12, 250.0, 0.0, dummy, dummy,
                              dummy, status.ctypes.data_as(_vp), None)
    assert rc == 1  # WT_ERR_ARG
