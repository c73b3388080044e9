"""The 3-D plane-wave expansion pinned by the mathematics.

The reference's expansion (planewvexp.f90, declared at interface.f90:43-47)
is missing, so the build defines it (oracle/rhs_oracle.py
``planewave_expand``).  A Thomson-Haskell field of a homogeneous
free-surface half-space (incident wave plus its free-surface reflections) is
an exact solution of the elastic equations, so on rows of the assembled
operator away from the PML and the surface, ||A u|| / ||omega^2 rho u|| must
be O((k dx)^2): it falls about 4x when dx halves at fixed frequency.  A
flipped horizontal phase sign, swapped horizontal polarisation or flipped
vertical component leaves an O(1) residual that does not fall.

The reference's interior stencil puts lambda + 2 mu on d^2/dx^2 and mu on
d^2/dy^2 in the y-component (v) row -- the reverse of isotropic elasticity,
whose v row has lambda + 2 mu on d^2/dy^2 (checked below on the compiled
reference and on the oracle).  The assembly reproduces the reference bit
for bit, so a plane wave is a discrete near-solution only where that swap
does not act: v = 0, or cos^2(phi) = sin^2(phi).  The azimuths tested are
0, 135 and 180 degrees.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import rhs_oracle as RO  # noqa: E402
import solve_oracle as SO  # noqa: E402

L, H, FREQ = 8000.0, 6000.0, 0.5
VP, VS, RHO = 6000.0, 6000.0 / np.sqrt(3.0), 2700.0
OMEGA = 2 * np.pi * FREQ
DXS = (500.0, 250.0, 125.0)
LAYERS = dict(top=[0.0, 1e9], vp=[VP] * 2, vs=[VS] * 2, rho=[RHO] * 2, vpd=[VP] * 2, vsd=[VS] * 2)


def grid(dx):
    return int(L / dx), int(L / dx), int(H / dx)


_OPS = {}


def operator(dx):
    """Oracle assembly of the homogeneous half-space (free surface on top,
    a thin 2-interval PML elsewhere) and the mask of the rows tested."""
    if dx not in _OPS:
        nx, ny, nz = grid(dx)
        sh = (ny, nx, nz)
        lam, mu = O.lame(np.full(sh, VP), np.full(sh, VS), np.full(sh, RHO), 64)
        ipml, pmlf = O.pml_params(dx, 2 * dx, 1e-3, 64)
        rc, vals = O.assemble(nx, ny, nz, dx, OMEGA, lam, mu, np.full(nx * ny * nz, RHO), 2 * dx, pmlf, ipml,
                              prec=64)
        assert rc == 0
        A = SO.csr(*O.graph(nx, ny, nz), vals)
        iy, ix, iz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
        m = (ix >= ipml) & (ix < nx - ipml) & (iy >= ipml) & (iy < ny - ipml) & (iz >= 2) & (iz < nz - ipml)
        _OPS[dx] = (A, np.repeat(m.ravel(), 3))
    return _OPS[dx]


def profile(dx, ain, is_s):
    nz = grid(dx)[2]
    rc, u, w, kh = O.haskgrn(LAYERS, ain, is_s, OMEGA, nz, dx, 0.0, 64)
    assert rc == 0
    return u, w, kh


def interior_residual(dx, U):
    A, m = operator(dx)
    r = A @ U
    return float(np.linalg.norm(r[m]) / np.linalg.norm(OMEGA ** 2 * RHO * U[m]))


def field(dx, ain, is_s, phi, mistake=None):
    """The oracle's expansion, or one with a planted mistake."""
    nx, ny, nz = grid(dx)
    u, w, kh = profile(dx, ain, is_s)
    if mistake == "phase_sign":
        kh = -kh
    U = RO.planewave_expand(nx, ny, nz, dx, u, w, [kh], [phi])[:, 0].reshape(-1, 3).copy()
    if mistake == "swap_horizontal":
        U[:, [0, 1]] = U[:, [1, 0]]
    if mistake == "vertical_sign":
        U[:, 2] *= -1
    return U.ravel()


@pytest.mark.parametrize("phi_deg", [0.0, 135.0, 180.0])
@pytest.mark.parametrize("is_s", [0, 1])
def test_expansion_is_a_second_order_discrete_solution(phi_deg, is_s):
    ain, phi = np.deg2rad(20.0), np.deg2rad(phi_deg)
    r = [interior_residual(dx, field(dx, ain, is_s, phi)) for dx in DXS]
    assert r[-1] <= 1e-3, r
    for a, b in zip(r, r[1:]):
        assert 3.0 <= a / b <= 6.0, r  # O(dx^2)


@pytest.mark.parametrize("mistake,phi_deg", [("phase_sign", 0.0), ("phase_sign", 135.0), ("swap_horizontal", 0.0),
                                             ("swap_horizontal", 135.0), ("vertical_sign", 0.0)])
@pytest.mark.parametrize("is_s", [0, 1])
def test_planted_mistakes_do_not_converge(mistake, phi_deg, is_s):
    ain, phi = np.deg2rad(20.0), np.deg2rad(phi_deg)
    r = [interior_residual(dx, field(dx, ain, is_s, phi, mistake)) for dx in DXS]
    good = interior_residual(DXS[-1], field(DXS[-1], ain, is_s, phi))
    assert min(r) >= 0.1 and r[-1] >= 100 * good, (r, good)


def _quadratic_rows(A, n, dx):
    iy, ix, iz = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x, y = ix.ravel() * dx, iy.ravel() * dx
    o = 0 * x
    c = np.nonzero(((iy == n // 2) & (ix == n // 2) & (iz == n // 2)).ravel())[0][0]
    out = {}
    for name, f in (("x2u", (x * x, o, o)), ("y2u", (y * y, o, o)), ("x2v", (o, x * x, o)), ("y2v", (o, y * y, o)),
                    ("xyu", (x * y, o, o))):
        u = np.stack(f, 1).ravel()
        out[name] = (A @ u)[3 * c:3 * c + 3].real
    return out


def _check_v_row(rows):
    lam = VP ** 2 * RHO - 2 * VS ** 2 * RHO
    mu = VS ** 2 * RHO
    tol = 1e-9 * (lam + 2 * mu)
    # u row: lambda + 2 mu on d2/dx2, mu on d2/dy2 (isotropic elasticity)
    assert abs(rows["x2u"][0] - 2 * (lam + 2 * mu)) <= tol and abs(rows["y2u"][0] - 2 * mu) <= tol
    # v row: the reference's placement, lambda + 2 mu on d2/dx2
    assert abs(rows["x2v"][1] - 2 * (lam + 2 * mu)) <= tol and abs(rows["y2v"][1] - 2 * mu) <= tol
    # mixed term: lambda + mu
    assert abs(rows["xyu"][1] - (lam + mu)) <= tol


def test_v_row_coefficients_oracle():
    n, dx = 12, 250.0
    sh = (n, n, n)
    lam, mu = O.lame(np.full(sh, VP), np.full(sh, VS), np.full(sh, RHO), 64)
    ipml, pmlf = O.pml_params(dx, 2 * dx, 1e-3, 64)
    rc, vals = O.assemble(n, n, n, dx, 0.0, lam, mu, np.full(n ** 3, RHO), 2 * dx, pmlf, ipml, prec=64)
    _check_v_row(_quadratic_rows(SO.csr(*O.graph(n, n, n), vals), n, dx))


@pytest.mark.skipif(not O.have_ref(), reason="reference build oracle/_ref absent")
def test_v_row_coefficients_reference():
    n, dx = 12, 250.0
    sh = (n, n, n)
    rc, ir, jc, a = O.ref_assemble(n, n, n, dx, 0.0, np.full(sh, VP), np.full(sh, VS), np.full(sh, RHO), 2 * dx,
                                   1e-3, prec=64)
    assert rc == 0
    _check_v_row(_quadratic_rows(SO.csr(ir, jc, a), n, dx))
