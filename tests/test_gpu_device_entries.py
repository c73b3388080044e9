"""The device-array entry points against their host-array counterparts, and
the small utility entries, through the C ABI: wt_graph_dev (SREL3D +
SPARSE_ADJ2CRS, srgraphs.f:638-734, sparse.f:13-49) against wt_graph,
wt_assemble_dev (ASMBLE, asmble.f90:1-636) against wt_assemble with its
negative-PML-depth flag (pml3d.f:219-223), wt_solver_planewaves against its
formula and against the host generator the benchmark uses, wt_device_count."""
import ctypes

import numpy as np
import pytest

import wtomo3d_amd as W
from wtomo3d_amd import media
from wtomo3d_amd.rhs import DeviceBuffer
from wtomo3d_amd.solver import Solver
from solve_cases import problem

pytestmark = pytest.mark.gpu
_vp = ctypes.c_void_p
P = lambda a: a.ctypes.data_as(_vp)  # noqa: E731


@pytest.mark.parametrize("dims,base", [((5, 4, 6), 1), ((9, 7, 3), 0), ((2, 2, 2), 1)])
def test_graph_dev_equals_graph(dims, base):
    L = W.lib()
    nx, ny, nz = dims
    n = 3 * nx * ny * nz
    nnz = 9 * (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2)
    rh, ch = np.zeros(n + 1, np.int32), np.zeros(nnz, np.int32)
    assert L.wt_graph(nx, ny, nz, base, P(rh), P(ch)) == 0
    R, C = DeviceBuffer((n + 1) * 4), DeviceBuffer(nnz * 4)
    assert L.wt_graph_dev(nx, ny, nz, base, R.ptr, C.ptr, None) == 0
    assert L.wt_dev_sync() == 0
    assert np.array_equal(R.to_host(np.int32, n + 1), rh)
    assert np.array_equal(C.to_host(np.int32, nnz), ch)
    assert rh[0] == base and rh[-1] == nnz + base
    # a degenerate grid is refused before any launch
    assert L.wt_graph_dev(0, ny, nz, base, R.ptr, C.ptr, None) == W._lib.WT_ERR_ARG


def _model(nx, ny, nz, prec):
    prm, lam, mu, rho, ir, jc, vals = problem(nx, ny, nz, seed=21)
    ct, ft = (np.complex64, np.float32) if prec == 32 else (np.complex128, np.float64)
    return prm, lam.astype(ct), mu.astype(ct), np.asarray(rho, ft)


@pytest.mark.parametrize("prec", [64, 32])
def test_assemble_dev_equals_assemble(prec):
    L = W.lib()
    nx, ny, nz = 9, 8, 10
    prm, lam, mu, rho = _model(nx, ny, nz, prec)
    g = W._lib.WtGridParams(nx, ny, nz, prm["ipml"], 1, 1, prm["dx"], prm["pmld"], prm["pmlf"], prm["omega"], 0.0)
    nnz = 9 * (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2)
    ct = np.complex64 if prec == 32 else np.complex128
    vh = np.zeros(nnz, ct)
    assert L.wt_assemble(ctypes.byref(g), prec, P(lam), P(mu), P(rho), P(vh)) == 0
    dl, dm, dr = DeviceBuffer.from_host(lam), DeviceBuffer.from_host(mu), DeviceBuffer.from_host(rho)
    dv, de = DeviceBuffer(vh.nbytes), DeviceBuffer.from_host(np.zeros(1, np.int32))
    assert L.wt_assemble_dev(ctypes.byref(g), prec, dl.ptr, dm.ptr, dr.ptr, dv.ptr, de.ptr, None) == 0
    assert L.wt_dev_sync() == 0
    assert dv.to_host(ct, nnz).tobytes() == vh.tobytes()
    assert de.to_host(np.int32, 1)[0] == 0
    # no error word: the same values
    dv2 = DeviceBuffer(vh.nbytes)
    assert L.wt_assemble_dev(ctypes.byref(g), prec, dl.ptr, dm.ptr, dr.ptr, dv2.ptr, None, None) == 0
    assert L.wt_dev_sync() == 0
    assert dv2.to_host(ct, nnz).tobytes() == vh.tobytes()


def test_assemble_dev_flags_negative_pml_depth():
    """A PML thicker (ipml) than its depth pmld allows gives nodes a negative
    damping depth: the device error word gets bit 1 and the host entry
    returns WT_ERR_ARG, as the reference stops there (pml3d.f:219-223)."""
    L = W.lib()
    nx, ny, nz = 12, 11, 10
    prm, lam, mu, rho = _model(nx, ny, nz, 64)
    pmld = 2.0 * prm["dx"]
    ipml_ok = int(pmld / prm["dx"]) + 1
    nnz = 9 * (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2)
    dl, dm, dr = DeviceBuffer.from_host(lam), DeviceBuffer.from_host(mu), DeviceBuffer.from_host(rho)
    dv = DeviceBuffer(nnz * 16)
    for ipml, bad in ((ipml_ok, False), (ipml_ok + 2, True)):
        g = W._lib.WtGridParams(nx, ny, nz, ipml, 1, 1, prm["dx"], pmld, prm["pmlf"], prm["omega"], 0.0)
        de = DeviceBuffer.from_host(np.zeros(1, np.int32))
        assert L.wt_assemble_dev(ctypes.byref(g), 64, dl.ptr, dm.ptr, dr.ptr, dv.ptr, de.ptr, None) == 0
        assert L.wt_dev_sync() == 0
        assert de.to_host(np.int32, 1)[0] == (1 if bad else 0)
        vh = np.zeros(nnz, np.complex128)
        rc = L.wt_assemble(ctypes.byref(g), 64, P(lam), P(mu), P(rho), P(vh))
        assert rc == (W._lib.WT_ERR_ARG if bad else 0)


def test_planewaves_formula_and_host_generator():
    """wt_solver_planewaves: u_j(r) = d_j exp(-i k d_j.r) on the grid in the
    reference's unknown order (u, v, w per node; z fastest, then x, then y),
    and the same fields as media.plane_wave_fields (k = omega / vp)."""
    L = W.lib()
    nx, ny, nz, nrhs, dx = 7, 5, 6, 5, 200.0
    omega, vp = 2 * np.pi * 0.3, 5200.0
    Uh = media.plane_wave_fields(nx, ny, nz, dx, omega, vp, nrhs, seed=4)
    n = 3 * nx * ny * nz
    # node 0 sits at r = 0, so its rows are the directions the host generator drew
    dirs = np.stack([Uh[0], Uh[1], Uh[2]], axis=1).real.copy()
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=1e-14)
    with Solver(nx, ny, nz) as s:
        X = DeviceBuffer(n * nrhs * 16)
        d = np.ascontiguousarray(dirs, np.float64)
        assert L.wt_solver_planewaves(s._h, nrhs, P(d), omega / vp, dx, X.ptr) == 0
        Ud = X.to_host(np.complex128, (n, nrhs))
    iy, ix, iz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    r = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1) * dx
    ph = np.exp(-1j * (omega / vp) * (r @ d.T))  # (nodes, nrhs)
    Uf = np.empty((n, nrhs), np.complex128)
    for c in range(3):
        Uf[c::3] = d[:, c][None, :] * ph
    assert np.max(np.abs(Ud - Uf)) <= 1e-13
    assert np.max(np.abs(Ud - Uh)) <= 1e-13
    # a null direction array is refused
    with Solver(nx, ny, nz) as s:
        assert L.wt_solver_planewaves(s._h, nrhs, None, 1.0, dx, X.ptr) == W._lib.WT_ERR_ARG


def test_device_count():
    assert W.lib().wt_device_count() >= 1
