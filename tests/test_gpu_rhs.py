"""GPU plane-wave right-hand sides through the C ABI: Thomson-Haskell
profiles (wt_haskgrn) against the reference's golden profiles and the
oracle, the 3-D expansion (wt_planewave_layered) against the oracle's and as a
discrete solution of the device operator, and
the scattered-field source (wt_solver_scattered_source) against the
reference's HETFLDSRC fixture and the oracle.

Tolerances: the device evaluates the same operations in the same order;
only the double cos/sin/exp of the device math library differ from the
host's in the last bits, and the single-precision sine of the incidence
angle is the rounded double sine.  FP64 profiles: 1e-10 of the largest
value; FP32 profiles (complex64 output): 1e-5 of the largest value."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from helpers import load, ref_params
from test_rhs_oracle import hk_golden, LAYER_KEYS, _golden_scatter
from wtomo3d_amd import media, rhs
from wtomo3d_amd._lib import grid_params
from wtomo3d_amd.solver import Solver

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import rhs_oracle as RO  # noqa: E402

pytestmark = pytest.mark.gpu
TOL = {32: 1e-5, 64: 1e-10}


def _close(a, ref, prec):
    scale = max(np.nanmax(np.abs(ref)), 1e-30)
    return np.nanmax(np.abs(np.asarray(a, np.complex128) - ref)) <= TOL[prec] * scale


@pytest.mark.parametrize("prec", [32, 64])
def test_haskgrn_golden(prec):
    ft = np.float32 if prec == 32 else np.float64
    for name, lay, ain, is_s, om, nz, dz, zorig, g in hk_golden():
        lay = {k: np.asarray(lay[k], ft) for k in LAYER_KEYS}
        err = int(g["%s__ierr%d" % (name, prec)])
        if err:
            with pytest.raises(RuntimeError):
                rhs.profiles(lay, ft(ain), is_s, ft(om), nz, ft(dz), ft(zorig), prec)
            continue
        pr = rhs.profiles(lay, ft(ain), is_s, ft(om), nz, ft(dz), ft(zorig), prec)
        U, W, kh = pr.host()
        print(name, prec, flush=True)
        assert _close(U[0], g["%s__u%d" % (name, prec)], prec), name
        assert _close(W[0], g["%s__w%d" % (name, prec)], prec), name
        _, _, _, pic = O.haskgrn(lay, ft(ain), is_s, ft(om), nz, ft(dz), ft(zorig), prec)
        assert kh[0] == pytest.approx(pic, rel=1e-14)


@pytest.mark.parametrize("prec", [32, 64])
def test_haskgrn_many_sources_vs_oracle(prec):
    """48 sources (P and S, 5-75 degrees) on the layered background of the
    96x96x64 configuration, one launch: every profile matches the oracle."""
    ft = np.float32 if prec == 32 else np.float64
    lay = {k: np.asarray(v, ft) for k, v in media.layer_stack(96, 96, 64, 200.0, 1003).items()}
    rng = np.random.default_rng(9)
    ns, nz, om, dz = 48, 64, ft(2 * np.pi * 0.3), ft(200.0)
    ain = rng.uniform(np.deg2rad(5.0), np.deg2rad(75.0), ns).astype(ft)
    iss = (rng.random(ns) < 0.5).astype(np.int32)
    U, W, kh = rhs.profiles(lay, ain, iss, om, nz, dz, 0.0, prec).host()
    for j in range(ns):
        rc, u, w, pic = O.haskgrn(lay, ain[j], iss[j], om, nz, dz, ft(0.0), prec)
        assert rc == 0
        assert _close(U[j], u, prec) and _close(W[j], w, prec), j
        assert kh[j] == pytest.approx(pic, rel=1e-14)


def test_layered_plane_wave_expansion():
    """The device expansion against the oracle's (oracle/rhs_oracle.py,
    itself pinned by the convergence test of test_planewave_consistency)."""
    nx, ny, nz, dx = 7, 6, 9, 250.0
    lay = media.layer_stack(nx, ny, nz, dx, 5)
    ain = np.deg2rad([12.0, 30.0, 44.0])
    phi = np.deg2rad([0.0, 135.0, 290.0])
    pr = rhs.profiles(lay, ain, [0, 1, 0], 2 * np.pi * 0.25, nz, dx, 0.0, 64)
    X = rhs.background_fields(nx, ny, nz, dx, pr, phi).to_host(np.complex128, (3 * nx * ny * nz, 3))
    U, W, kh = pr.host()
    ref = RO.planewave_expand(nx, ny, nz, dx, U, W, kh, phi)
    assert np.abs(X - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("phi_deg,is_s", [(0.0, 0), (135.0, 1)])
def test_expansion_converges_on_the_device_operator(phi_deg, is_s):
    """The device profiles, expansion and FP64 assembly together: the
    homogeneous half-space's Haskell field is a second-order discrete
    solution on the interior rows (the property the oracle's expansion is
    pinned by), falling about 4x per halving of dx."""
    import test_planewave_consistency as PC
    r = []
    for dx in PC.DXS:
        nx, ny, nz = PC.grid(dx)
        pr = rhs.profiles(PC.LAYERS, [np.deg2rad(20.0)], [is_s], PC.OMEGA, nz, dx, 0.0, 64)
        s = Solver(nx, ny, nz)
        sh = (ny, nx, nz)
        s.set_model(*media.lame(np.full(sh, PC.VP), np.full(sh, PC.VS), np.full(sh, PC.RHO)))
        s.assemble(grid_params(nx, ny, nz, dx, PC.OMEGA, pmld=2 * dx))
        X = rhs.background_fields(nx, ny, nz, dx, pr, [np.deg2rad(phi_deg)]).to_host(np.complex128,
                                                                                    (3 * nx * ny * nz, 1))
        AX = s.apply(X)[:, 0]
        s.close()
        _, m = PC.operator(dx)
        r.append(float(np.linalg.norm(AX[m]) / np.linalg.norm(PC.OMEGA ** 2 * PC.RHO * X[m, 0])))
        print("dx %.0f interior residual %.3e" % (dx, r[-1]), flush=True)
    assert r[-1] <= 1e-3
    for a, b in zip(r, r[1:]):
        assert 3.0 <= a / b <= 6.0, r


def _golden_solver(tau=999.999):
    g, nx, ny, nz, dx, freq, pmld, pmlr, srcfn = _golden_scatter(64)
    meta = dict(dx=dx, freq=freq, pmld=pmld, pmlr=pmlr, tau=tau)
    dxp, om, pmldp, pmlf, ipml = ref_params(meta, 64)
    s = Solver(nx, ny, nz)
    lam, mu, rho = media.lame(g["vp"], g["vs"], g["rho"])
    s.set_model(lam, mu, rho)
    p = grid_params(nx, ny, nz, dxp, om, pmld=pmldp)
    p.pmlf, p.ipml = pmlf, ipml
    s.assemble(p)
    u = RO.nodes_to_unknowns(g["ub"], g["vb"], g["wb"])
    return g, s, u, srcfn


def test_scattered_source_matches_reference_fixture():
    g, s, u, srcfn = _golden_solver()
    lb, mb, rb = media.lame(g["vpb"], g["vsb"], g["rhob"])
    sf = s.scattered_source(lb, mb, rb, u[:, None], srcfn)[:, 0]
    ref = g["sfld64"]
    assert np.abs(sf - ref).max() <= 1e-10 * np.abs(ref).max()
    s.close()


@pytest.mark.skipif(not O.have_ref(), reason="reference build absent")
@pytest.mark.parametrize("tau", [3.0, 0.7])
def test_scattered_source_complex_frequency_vs_reference(tau):
    """With the anti-aliasing damping (omega = 2 pi f + i / tau, elwave.f90:164)
    the handle assembled at the complex frequency gives the reference's
    HETFLDSRC at that frequency."""
    g, s, u, srcfn = _golden_solver(tau)
    g2, nx, ny, nz, dx, freq, pmld, pmlr, _ = _golden_scatter(64)
    lb, mb, rb = media.lame(g["vpb"], g["vsb"], g["rhob"])
    sf = s.scattered_source(lb, mb, rb, u[:, None], srcfn)[:, 0]
    s.close()
    ierr, ref = O.ref_hetfldsrc(nx, ny, nz, dx, freq, g["vp"], g["vs"], g["rho"], g["vpb"], g["vsb"], g["rhob"],
                                g["ub"], g["vb"], g["wb"], srcfn, pmld, pmlr, tau=tau, prec=64)
    assert ierr == 0
    assert np.abs(sf - ref).max() <= 1e-10 * np.abs(ref).max()


def test_scattered_source_invariants_and_wide_blocks():
    """Background == model gives zero; 70 columns (two column groups) give
    the per-column results, each scaled by its srcfn."""
    g, s, u, srcfn = _golden_solver()
    lam, mu, rho = media.lame(g["vp"], g["vs"], g["rho"])
    z = s.scattered_source(lam, mu, rho, u[:, None], srcfn)
    assert not np.any(z)
    lb, mb, rb = media.lame(g["vpb"], g["vsb"], g["rhob"])
    rng = np.random.default_rng(3)
    B = u[:, None] * (1.0 + 0.1 * rng.standard_normal(70))[None, :]
    f = rng.standard_normal(70) + 1j * rng.standard_normal(70)
    S = s.scattered_source(lb, mb, rb, B, f)
    one = s.scattered_source(lb, mb, rb, u[:, None], None)[:, 0]
    for j in (0, 33, 63, 64, 69):
        ref = f[j] * (B[0, j] / u[0]) * one
        assert np.abs(S[:, j] - ref).max() <= 1e-12 * np.abs(ref).max()
    s.close()


def test_scattered_source_without_a_handle():
    """wt_scattered_source_grid (both operators assembled in scratch, host
    model arrays) equals the handle's source bit for bit, writes into a
    column range of a wider block (row strides), and matches the reference's
    HETFLDSRC fixture with its srcfn."""
    import ctypes
    from wtomo3d_amd import lib
    from wtomo3d_amd.rhs import DeviceBuffer
    g, s, u, srcfn = _golden_solver()
    lam, mu, rho = media.lame(g["vp"], g["vs"], g["rho"])
    lb, mb, rb = media.lame(g["vpb"], g["vsb"], g["rhob"])
    rng = np.random.default_rng(5)
    nr, wide, c0 = 5, 9, 3
    B = np.ascontiguousarray(u[:, None] * (1.0 + 0.1 * rng.standard_normal(nr))[None, :])
    ref = s.scattered_source(lb, mb, rb, B, None)
    n = s.n
    s.close()
    _, nx, ny, nz, dx, freq, pmld, pmlr, _ = _golden_scatter(64)
    dxp, om, pmldp, pmlf, ipml = ref_params(dict(dx=dx, freq=freq, pmld=pmld, pmlr=pmlr), 64)
    p = grid_params(nx, ny, nz, dxp, om, pmld=pmldp)
    p.pmlf, p.ipml = pmlf, ipml
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    dU = DeviceBuffer.from_host(B)
    dO = DeviceBuffer.from_host(np.zeros((n, wide), np.complex128))
    dst = ctypes.c_void_p(dO.ptr.value + c0 * 16)
    assert lib().wt_scattered_source_grid(ctypes.byref(p), P(lam), P(mu), P(rho), P(lb), P(mb), P(rb), 0, nr,
                                          dU.ptr, nr, None, dst, wide, None) == 0
    O = dO.to_host(np.complex128, (n, wide))
    assert np.array_equal(O[:, c0:c0 + nr], ref)
    assert not np.any(O[:, :c0]) and not np.any(O[:, c0 + nr:])
    du = DeviceBuffer.from_host(np.ascontiguousarray(u))
    f = np.ascontiguousarray([srcfn], np.complex128)
    assert lib().wt_scattered_source_grid(ctypes.byref(p), P(lam), P(mu), P(rho), P(lb), P(mb), P(rb), 0, 1,
                                          du.ptr, 1, P(f), dO.ptr, 1, None) == 0
    sf = dO.to_host(np.complex128, (n,))
    assert np.abs(sf - g["sfld64"]).max() <= 1e-10 * np.abs(g["sfld64"]).max()


def test_total_field_from_scattered_source():
    """Plane waves in the layered background, scattered by a perturbed model:
    solving A u_h = (A_bg - A) u_bg gives a total field u_h + u_bg with
    A (u_h + u_bg) = A_bg u_bg (hetfldsrc.f90 and makeuh.f:36-71)."""
    nx, ny, nz, dx = 10, 9, 12, 300.0
    om = 2 * np.pi * 0.25
    lay = media.layer_stack(nx, ny, nz, dx, 21)
    vp1, vs1, rho1 = media.layered(nx, ny, nz, dx, 21)
    vpb = np.broadcast_to(vp1[None, None, :], (ny, nx, nz))
    vsb = np.broadcast_to(vs1[None, None, :], (ny, nx, nz))
    rhob = np.broadcast_to(rho1[None, None, :], (ny, nx, nz))
    d = media.gaussian_field((ny, nx, nz), dx, 900.0, 0.05, 4)
    s = Solver(nx, ny, nz)
    lam, mu, rho = media.lame(vpb * (1 + d), vsb * (1 + d), rhob * (1 + 0.3 * d))
    s.set_model(lam, mu, rho)
    s.assemble(grid_params(nx, ny, nz, dx, om, pmld=4 * dx))
    pr = rhs.profiles(lay, np.deg2rad([15.0, 28.0]), [0, 1], om, nz, dx, 0.0, 64)
    ubg = rhs.background_fields(nx, ny, nz, dx, pr, np.deg2rad([30.0, 200.0]))
    lb, mb, rb = media.lame(vpb, vsb, rhob)
    sf = s.scattered_source(lb, mb, rb, ubg, 1.0).to_host(np.complex128, (s.n, 2))
    Ubg = ubg.to_host(np.complex128, (s.n, 2))
    s.factor(64)
    uh, st = s.solve(sf, prec=64, tol=1e-12, max_iter=3)
    ut = uh + Ubg
    lhs = s.apply(ut)
    # A_bg u_bg with the background operator
    sb = Solver(nx, ny, nz)
    sb.set_model(lb, mb, rb)
    sb.assemble(grid_params(nx, ny, nz, dx, om, pmld=4 * dx))
    rhs_bg = sb.apply(Ubg)
    for j in range(2):
        assert np.linalg.norm(lhs[:, j] - rhs_bg[:, j]) <= 1e-9 * np.linalg.norm(rhs_bg[:, j])
    s.close()
    sb.close()
