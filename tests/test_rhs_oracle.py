"""The plane-wave right-hand-side oracle (CPU): Thomson-Haskell profiles
(oracle/hask_oracle.c) and the scattered-field source (oracle/rhs_oracle.py)
against the compiled reference, its golden fixtures and closed forms."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from helpers import GOLDEN, load

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import rhs_oracle as RO  # noqa: E402

needs_ref = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built")
LAYER_KEYS = ("top", "vp", "vs", "rho", "vpd", "vsd")


def hk_golden():
    g = load("haskgrn_cases")
    out = []
    for name in g["names"]:
        name = str(name)
        lay = {k: g["%s__%s" % (name, k)] for k in LAYER_KEYS}
        ain, is_s, om, nz, dz, zorig = g[name + "__scalars"]
        out.append((name, lay, float(ain), int(is_s), float(om), int(nz), float(dz), float(zorig), g))
    return out


def _cast(lay, prec, scal):
    ft = np.float32 if prec == 32 else np.float64
    return {k: np.asarray(lay[k], ft) for k in LAYER_KEYS}, [float(ft(v)) for v in scal]


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("case", range(6))
def test_haskgrn_golden(case, prec):
    """Oracle profiles equal the reference's (golden fixtures) bit for bit."""
    name, lay, ain, is_s, om, nz, dz, zorig, g = hk_golden()[case]
    lay, (ain, om, dz, zorig) = _cast(lay, prec, (ain, om, dz, zorig))
    rc, u, w, _ = O.haskgrn(lay, ain, is_s, om, nz, dz, zorig, prec)
    assert rc == int(g["%s__ierr%d" % (name, prec)])
    np.testing.assert_array_equal(u, g["%s__u%d" % (name, prec)])
    np.testing.assert_array_equal(w, g["%s__w%d" % (name, prec)])


@needs_ref
@pytest.mark.parametrize("prec", [32, 64])
def test_haskgrn_random_models_live_reference(prec):
    """Seeded random layered models (thin layers the grid skips, fluid
    layers, low-velocity zones, evanescent waves): oracle == reference."""
    rng = np.random.default_rng(77 + prec)
    for _ in range(120):
        nl = int(rng.integers(2, 9))
        vp = rng.uniform(1500, 8500, nl)
        if rng.random() < 0.5:
            vp = np.sort(vp)
        vs = vp / rng.uniform(1.6, 2.2, nl)
        if nl > 2 and rng.random() < 0.2:
            vs[int(rng.integers(1, nl - 1))] = 0.0
        th = rng.choice([rng.uniform(50, 400), rng.uniform(300, 6000)], nl - 1)
        lay = dict(vp=vp, vs=vs, rho=rng.uniform(1000, 3300, nl), top=np.concatenate([[0.0], np.cumsum(th)]),
                   vpd=vp * rng.uniform(0.99, 1.0, nl), vsd=vs * rng.uniform(0.99, 1.0, nl))
        args = (float(rng.uniform(0.05, 1.2)), int(rng.random() < 0.5), float(2 * np.pi * rng.uniform(0.05, 0.5)),
                int(rng.integers(1, 80)), float(rng.uniform(100, 800)), float(rng.choice([0.0, 10.0, -900.0])))
        ft = np.float32 if prec == 32 else np.float64
        lay = {k: np.asarray(v, ft) for k, v in lay.items()}
        rc, u, w, _ = O.haskgrn(lay, *args, prec=prec)
        ie, ur, wr = O.ref_haskgrn(lay, *args, prec=prec)
        assert rc == ie
        np.testing.assert_array_equal(u, ur)
        np.testing.assert_array_equal(w, wr)


def _halfspace_exact(alpha, beta, rho, p, omega, z, wave):
    """Free-surface response of a homogeneous half-space to a unit plane P or
    SV wave from below (traction-free surface solved for the reflected P and
    SV amplitudes), exp(i w (t - p x - q z)), z down.  Returns (u_x, u_z)."""
    lam, mu = rho * (alpha ** 2 - 2 * beta ** 2), rho * beta ** 2
    ea, eb = np.sqrt(1 / alpha ** 2 - p ** 2 + 0j), np.sqrt(1 / beta ** 2 - p ** 2 + 0j)

    def trac(d, q):
        return np.array([lam * (p * d[0] + q * d[1]) + 2 * mu * q * d[1], mu * (q * d[0] + p * d[1])])

    if wave == "P":
        qi, di = -ea, np.array([p * alpha, -ea * alpha])
    else:
        qi, di = -eb, np.array([-eb * beta, -p * beta])
    dP, dS = np.array([p * alpha, ea * alpha]), np.array([eb * beta, -p * beta])
    rp, rs = np.linalg.solve(np.stack([trac(dP, ea), trac(dS, eb)], axis=1), -trac(di, qi))
    e = lambda q: np.exp(-1j * omega * q * z)  # noqa: E731
    ux = di[0] * e(qi) + rp * dP[0] * e(ea) + rs * dS[0] * e(eb)
    uz = di[1] * e(qi) + rp * dP[1] * e(ea) + rs * dS[1] * e(eb)
    return ux, uz, -qi


@pytest.mark.parametrize("wave,ain", [("P", 0.05), ("P", 0.3), ("P", 0.8), ("S", 0.2), ("S", 0.5)])
def test_haskgrn_halfspace_closed_form(wave, ain):
    """Two identical layers are a homogeneous half-space: the profiles are
    incident + reflected P + reflected SV of a unit wave whose phase is
    referenced at the half-space top H (exp(-i w q H)); SV carries the
    opposite polarization sign to the one chosen here."""
    a, b, r, H = 6000.0, 3400.0, 2700.0, 3000.0
    lay = dict(vp=[a, a], vs=[b, b], rho=[r, r], top=[0.0, H], vpd=[a, a], vsd=[b, b])
    om, nz, dz = 2 * np.pi * 0.4, 40, 150.0
    rc, u, w, pic = O.haskgrn(lay, ain, wave == "S", om, nz, dz, 0.0, 64)
    assert rc == 0
    p = np.sin(ain) / (a if wave == "P" else b)
    assert pic == pytest.approx(om * p, rel=1e-14)
    ux, uz, qin = _halfspace_exact(a, b, r, p, om, dz * np.arange(nz), wave)
    ph = np.exp(-1j * om * qin * H) * (1.0 if wave == "P" else -1.0)
    scale = max(np.abs(ux).max(), np.abs(uz).max())
    assert np.abs(u - ph * ux).max() < 1e-11 * scale
    assert np.abs(w - ph * uz).max() < 1e-11 * scale


def test_haskgrn_above_top_is_an_error():
    lay = dict(vp=[5000.0, 7000.0], vs=[2900.0, 4000.0], rho=[2500.0, 3000.0], top=[0.0, 4000.0],
               vpd=[5000.0, 7000.0], vsd=[2900.0, 4000.0])
    rc, u, w, _ = O.haskgrn(lay, 0.3, 0, 2.0, 10, 200.0, -500.0, 64)
    assert rc == 1
    assert np.all(u[2:] == 0) and np.all(w[2:] == 0)


def _golden_scatter(prec):
    g = load("hetfldsrc_g877")
    nx, ny, nz = (int(v) for v in g["params"][:3])
    dx, freq, pmld, pmlr, sr, si = g["params"][3:]
    return g, nx, ny, nz, dx, freq, pmld, pmlr, complex(sr, si)


def _oracle_ops(g, nx, ny, nz, dx, freq, pmld, pmlr, prec, tau=999.999):
    from helpers import ref_params, lame_ref
    meta = dict(dx=dx, freq=freq, pmld=pmld, pmlr=pmlr, tau=tau)
    dxp, om, pmldp, pmlf, ipml = ref_params(meta, prec)
    ft = np.float32 if prec == 32 else np.float64
    ops = []
    for vp, vs, rho in ((g["vp"], g["vs"], g["rho"]), (g["vpb"], g["vsb"], g["rhob"])):
        vp, vs, rho = (np.asarray(x, ft) for x in (vp, vs, rho))
        lam, mu, rh = lame_ref(vp, vs, rho, prec)
        rc, a = O.assemble(nx, ny, nz, dxp, om, lam, mu, rh, pmldp, pmlf, ipml, prec=prec)
        assert rc == 0
        ops.append(a)
    ir, jc = O.graph(nx, ny, nz)
    return ir, jc, ops[0], ops[1]


def test_scattered_source_matches_reference_fp64():
    g, nx, ny, nz, dx, freq, pmld, pmlr, srcfn = _golden_scatter(64)
    ir, jc, a, abg = _oracle_ops(g, nx, ny, nz, dx, freq, pmld, pmlr, 64)
    u = RO.nodes_to_unknowns(g["ub"], g["vb"], g["wb"])
    sf = RO.scattered_source(ir, jc, a, abg, u, srcfn)
    ref = g["sfld64"]
    assert np.abs(sf - ref).max() <= 1e-11 * np.abs(ref).max()


def test_scattered_source_matches_reference_fp32():
    """The reference as compiled subtracts complex64 entries and sums in
    complex64: agreement to its rounding."""
    g, nx, ny, nz, dx, freq, pmld, pmlr, srcfn = _golden_scatter(32)
    ir, jc, a, abg = _oracle_ops(g, nx, ny, nz, dx, freq, pmld, pmlr, 32)
    u = RO.nodes_to_unknowns(g["ub"], g["vb"], g["wb"]).astype(np.complex64)
    sf = RO.scattered_source(ir, jc, a.astype(np.complex128), abg.astype(np.complex128), u, srcfn)
    ref = g["sfld32"]
    assert np.abs(sf - ref).max() <= 2e-5 * np.abs(ref).max()


def test_scattered_source_invariants():
    """Model == background gives exactly zero; linear in srcfn and u."""
    g, nx, ny, nz, dx, freq, pmld, pmlr, srcfn = _golden_scatter(64)
    ir, jc, a, abg = _oracle_ops(g, nx, ny, nz, dx, freq, pmld, pmlr, 64)
    u = RO.nodes_to_unknowns(g["ub"], g["vb"], g["wb"])
    assert not np.any(RO.scattered_source(ir, jc, abg, abg, u, srcfn))
    s1 = RO.scattered_source(ir, jc, a, abg, u, 1.0)
    np.testing.assert_allclose(RO.scattered_source(ir, jc, a, abg, u, srcfn), srcfn * s1, rtol=1e-14, atol=0)
    np.testing.assert_allclose(RO.scattered_source(ir, jc, a, abg, 2.5 * u, 1.0), 2.5 * s1, rtol=1e-14, atol=0)


@needs_ref
def test_scattered_source_background_equal_live_reference():
    """The reference itself returns exactly zero when model == background."""
    g, nx, ny, nz, dx, freq, pmld, pmlr, srcfn = _golden_scatter(64)
    ierr, sf = O.ref_hetfldsrc(nx, ny, nz, dx, freq, g["vpb"], g["vsb"], g["rhob"], g["vpb"], g["vsb"], g["rhob"],
                               g["ub"], g["vb"], g["wb"], srcfn, pmld, pmlr, prec=64)
    assert ierr == 0 and not np.any(sf)


@needs_ref
@pytest.mark.parametrize("tau", [3.0, 0.7])
def test_scattered_source_complex_frequency_live_reference(tau):
    """With the anti-aliasing damping set the reference forms both operators
    at omega = 2 pi f + i / tau (elwave.f90:164): the oracle's (A_bg - A) u_bg
    at that frequency against the compiled HETFLDSRC."""
    g, nx, ny, nz, dx, freq, pmld, pmlr, srcfn = _golden_scatter(64)
    ir, jc, a, abg = _oracle_ops(g, nx, ny, nz, dx, freq, pmld, pmlr, 64, tau=tau)
    u = RO.nodes_to_unknowns(g["ub"], g["vb"], g["wb"])
    sf = RO.scattered_source(ir, jc, a, abg, u, srcfn)
    ierr, ref = O.ref_hetfldsrc(nx, ny, nz, dx, freq, g["vp"], g["vs"], g["rho"], g["vpb"], g["vsb"], g["rhob"],
                                g["ub"], g["vb"], g["wb"], srcfn, pmld, pmlr, tau=tau, prec=64)
    assert ierr == 0
    assert np.abs(sf - ref).max() <= 1e-11 * np.abs(ref).max()
    # the damping changes the source
    assert np.abs(ref - g["sfld64"]).max() > 1e-6 * np.abs(ref).max()
