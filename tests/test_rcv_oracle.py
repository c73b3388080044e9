"""Receiver extraction oracle (oracle/rcv_oracle.py) against the reference's
MAKEUH (golden fixture and live build) and closed forms."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from helpers import load

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import rcv_oracle as RC  # noqa: E402

needs_ref = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built")


def golden_case():
    g = load("makeuh_g121110")
    nx, ny, nz = (int(v) for v in g["params"][:3])
    dx, freq = float(g["params"][3]), float(g["params"][4])
    ut = g["sfld"].reshape(-1, 3)
    fields = [ut[:, c].reshape(ny, nx, nz) + b for c, b in enumerate((g["ub"], g["vb"], g["wb"]))]
    lam = (g["vp"] ** 2 * g["rho"] - 2 * g["vs"] ** 2 * g["rho"]).astype(complex)
    mu = (g["vs"] ** 2 * g["rho"]).astype(complex)
    out = []
    for tag in g["sets"]:
        tag = str(tag)
        ireg, spread = g[tag + "__set"]
        st, hy = g[tag + "__stations"], g[tag + "__hyd"]
        rec = dict(x=st[0], y=st[1], z=st[2], spread=float(spread), ireg=int(ireg), fac=g[tag + "__fac"])
        hyd = dict(x=hy[0], y=hy[1], z=hy[2], spread=float(spread), ireg=int(ireg))
        out.append((tag, rec, hyd))
    return g, (nx, ny, nz, dx, freq), fields, lam, mu, out


@pytest.mark.parametrize("k", range(5))
def test_makeuh_matches_reference_fixture(k):
    g, (nx, ny, nz, dx, freq), (u, v, w), lam, mu, sets = golden_case()
    tag, rec, hyd = sets[k]
    res = np.concatenate(RC.makeuh(u, v, w, rec, hyd, lam, mu, dx, 2 * np.pi * freq, [1, 0, 0, 0, 0, 0]))
    ref64 = g[tag + "__out64"]
    assert np.abs(res - ref64).max() <= 1e-13 * np.abs(ref64).max()
    ref32 = g[tag + "__out32"]
    assert np.abs(res - ref32).max() <= 2e-5 * np.abs(ref32).max()


@needs_ref
def test_makeuh_live_reference_boundaries():
    """Stations near every face with a free surface on some faces only:
    mirrored and dropped taps both occur."""
    rng = np.random.default_rng(8)
    nx, ny, nz, dx, freq = 11, 12, 9, 50.0, 0.5
    vp = 3000 + 300 * rng.random((ny, nx, nz))
    vs, rho = vp / 1.9, 2200 + 50 * rng.random((ny, nx, nz))
    n = 3 * nx * ny * nz
    sfld = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    zero = np.zeros((ny, nx, nz), complex)
    for ireg, spread in ((3, 0.5), (2, 0.9)):
        rec = dict(x=[0.0, 20.0, (nx - 1) * dx, 270.0], y=[10.0, (ny - 1) * dx - 5, 300.0, 0.0],
                   z=[0.0, 10.0, (nz - 1) * dx, 200.0], spread=spread, ireg=ireg, fac=None)
        hyd = dict(x=[60.0], y=[80.0], z=[0.0], spread=spread, ireg=ireg)
        ie, *ref = O.ref_makeuh(nx, ny, nz, dx, freq, vp, vs, rho, sfld, zero, zero, zero, rec, hyd, prec=64)
        assert ie == 0
        ut = sfld.reshape(-1, 3)
        flds = [ut[:, c].reshape(ny, nx, nz) for c in range(3)]
        lam = (vp ** 2 * rho - 2 * vs ** 2 * rho).astype(complex)
        res = RC.makeuh(*flds, rec, hyd, lam, (vs ** 2 * rho).astype(complex), dx, 2 * np.pi * freq,
                        [1, 0, 0, 0, 0, 0])
        for a, b in zip(res, ref):
            assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def test_kaiser_sinc_on_node_is_exact():
    """A point receiver on a node reads that node (sinc zeros at integers)."""
    rng = np.random.default_rng(2)
    f = rng.standard_normal((9, 10, 11)) + 1j * rng.standard_normal((9, 10, 11))
    for ireg in (1, 2, 3, 4):
        val = RC.extract(f, 5 * 100.0, 4 * 100.0, 6 * 100.0, 0.5, ireg, 100.0, 100.0, 100.0, [0] * 6)
        assert val == pytest.approx(f[4, 5, 6], rel=1e-15, abs=1e-15)


def test_dirac_weights_reproduce_constants():
    """Unit-sum Dirac weights read a constant field exactly when the
    normalisation block (INT(4 s) + 1 nodes) is the extraction region
    (2 INT(2 s) + 1 nodes).  Otherwise (s = 0.8: 4 vs 3 nodes) the
    reference normalises over taps it never applies (srcreg.f:108-141,
    srcsetup.f:99-113) and the constant reads low."""
    f = np.full((12, 12, 12), 2.5 - 1.0j)
    for spread in (1.0, 1.2, 1.7):
        val = RC.extract(f, 512.3, 488.0, 530.9, spread, 1, 100.0, 100.0, 100.0, [0] * 6)
        assert val == pytest.approx(2.5 - 1.0j, rel=1e-14)
    val = RC.extract(f, 512.3, 488.0, 530.9, 0.8, 1, 100.0, 100.0, 100.0, [0] * 6)
    assert abs(val) < 0.95 * abs(2.5 - 1.0j)


def test_hydrophone_pressure_of_linear_field():
    """u = (a x, b y, c z): central differences are exact, so the pressure is
    (lambda + 2 mu / 3)(a + b + c) / (i omega)."""
    ny, nx, nz, dx = 8, 9, 10, 100.0
    iy, ix, iz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    a, b, c = 0.3 + 0.1j, -0.2, 0.05j
    u, v, w = a * ix * dx, b * iy * dx + 0j, c * iz * dx
    lam = np.full((ny, nx, nz), 3e10 + 1e8j)
    mu = np.full((ny, nx, nz), 2e10 + 0j)
    om = 2 * np.pi * 0.3
    hyd = dict(x=[400.0], y=[300.0], z=[500.0], spread=0.5, ireg=0)
    _, _, _, p = RC.makeuh(u, v, w, dict(x=[], y=[], z=[], spread=0.5, ireg=0), hyd, lam, mu, dx, om,
                           [1, 0, 0, 0, 0, 0])
    expect = (3e10 + 1e8j + 2 * 2e10 / 3) * (a + b + c) / (1j * om)
    assert p[0] == pytest.approx(expect, rel=1e-13)
