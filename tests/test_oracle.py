"""The assembly oracle pinned against the reference's outputs and the physics.

Golden vectors come from the reference's own Fortran (oracle/_ref, see
tests/golden/make_golden.py); the live comparison runs where _ref is built.
"""
import hashlib

import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as O
from helpers import cases, load, ref_params, lame_ref, rel_err_entries

CASE_NAMES = [k for k, v in cases().items() if isinstance(v, dict) and "seed" in v]


def test_closed_form_nnz():
    for nx, ny, nz in [(2, 2, 2), (3, 5, 4), (24, 24, 24), (64, 64, 48), (96, 96, 64), (80, 80, 80)]:
        n, nnz = O.graph_size(nx, ny, nz)
        assert n == 3 * nx * ny * nz
        assert nnz == 9 * (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2)


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 2, 4), (4, 3, 3), (5, 4, 3)])
def test_pattern_brute_force(dims):
    """Rows list every DOF of every node within Chebyshev distance 1, sorted."""
    nx, ny, nz = dims
    ir, jc = O.graph(nx, ny, nz)
    n = 3 * nx * ny * nz
    for row in range(n):
        g = row // 3
        iz, ix, iy = g % nz, (g // nz) % nx, g // (nx * nz)
        want = []
        for y in range(ny):
            for x in range(nx):
                for z in range(nz):
                    if max(abs(x - ix), abs(y - iy), abs(z - iz)) <= 1:
                        gg = z + x * nz + y * nx * nz
                        want += [3 * gg + 1, 3 * gg + 2, 3 * gg + 3]
        got = jc[ir[row] - 1: ir[row + 1] - 1]
        assert list(got) == sorted(want)


@pytest.mark.parametrize("name", CASE_NAMES)
def test_pattern_matches_reference(name):
    m = cases()[name]
    g = load(name)
    ir, jc = O.graph(m["nx"], m["ny"], m["nz"])
    assert np.array_equal(ir, g["irptr"]) and np.array_equal(jc, g["jcptr"])


def test_c1_pattern_digest():
    d = cases()["c1_pattern_sha256"]
    ir, jc = O.graph(24, 24, 24)
    assert hashlib.sha256(ir.tobytes()).hexdigest() == d["irptr"]
    assert hashlib.sha256(jc.tobytes()).hexdigest() == d["jcptr"]


def _oracle_case(name, prec):
    m = cases()[name]
    g = load(name)
    dx, om, pmld, pmlf, ipml = ref_params(m, prec)
    ft = np.float32 if prec == 32 else np.float64
    lam, mu, rho = lame_ref(g["vp"].astype(ft), g["vs"].astype(ft), g["rho"].astype(ft), prec)
    rc, a = O.assemble(m["nx"], m["ny"], m["nz"], dx, om, lam, mu, rho, pmld, pmlf, ipml,
                       fst=bool(m["fst"]), usemin=bool(m.get("usemin", 1)), prec=prec)
    assert rc == 0
    return a, g["vals%d" % prec]


@pytest.mark.parametrize("name", CASE_NAMES)
def test_values_f64_match_reference(name):
    a, ref = _oracle_case(name, 64)
    rel, zero_mismatch = rel_err_entries(a, ref)
    assert zero_mismatch == 0
    assert rel <= 1e-12


@pytest.mark.parametrize("name", CASE_NAMES)
def test_values_f32_match_reference(name):
    a, ref = _oracle_case(name, 32)
    rel, zero_mismatch = rel_err_entries(a, ref)
    assert zero_mismatch == 0
    assert rel <= 3e-5  # a few float32 ulps; the reference's complex division differs


def test_c1_sampled_rows():
    s = load("c1_rows")
    nx = ny = nz = 24
    ir, jc = O.graph(nx, ny, nz)
    idx = np.concatenate([np.arange(ir[r] - 1, ir[r + 1] - 1) for r in s["rows"]])
    meta = dict(dx=500.0, freq=0.2, pmld=5000.0, pmlr=1e-3)
    for prec, tol in ((64, 1e-12), (32, 3e-5)):
        ft = np.float32 if prec == 32 else np.float64
        vp = np.full((ny, nx, nz), 6000.0, ft)
        vs = (vp / ft(np.sqrt(3.0))).astype(ft)
        rho = np.full_like(vp, 2700.0)
        dx, om, pmld, pmlf, ipml = ref_params(meta, prec)
        lam, mu, rho1 = lame_ref(vp, vs, rho, prec)
        rc, a = O.assemble(nx, ny, nz, dx, om, lam, mu, rho1, pmld, pmlf, ipml, prec=prec)
        rel, zm = rel_err_entries(a[idx], s["vals%d" % prec])
        assert rc == 0 and zm == 0 and rel <= tol


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("prec", [32, 64])
def test_live_reference_random_model(prec):
    nx, ny, nz, dx = 11, 9, 10, 200.0
    rng = np.random.default_rng(77)
    ft = np.float32 if prec == 32 else np.float64
    vp = (4500 + 2500 * rng.random((ny, nx, nz))).astype(ft)
    vs = (vp / ft(1.73)).astype(ft)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(ft)
    ierr, ir, jc, ref = O.ref_assemble(nx, ny, nz, dx, 0.35, vp, vs, rho, 600.0, 1e-3, prec=prec)
    assert ierr == 0
    meta = dict(dx=dx, freq=0.35, pmld=600.0, pmlr=1e-3)
    dx_, om, pmld, pmlf, ipml = ref_params(meta, prec)
    lam, mu, rho1 = lame_ref(vp, vs, rho, prec)
    rc, a = O.assemble(nx, ny, nz, dx_, om, lam, mu, rho1, pmld, pmlf, ipml, prec=prec)
    rel, zm = rel_err_entries(a, ref)
    assert rc == 0 and zm == 0 and rel <= (1e-12 if prec == 64 else 3e-5)


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("tau", [4.0, 0.8])
def test_live_reference_complex_frequency(prec, tau):
    """The anti-aliasing damping: with tau set the reference assembles at the
    complex frequency omega = 2 pi f + i / tau (elwave.f90:164), which enters
    the mass terms and every PML stretch factor."""
    nx, ny, nz, dx = 12, 10, 9, 250.0
    rng = np.random.default_rng(78)
    ft = np.float32 if prec == 32 else np.float64
    vp = (4000 + 3000 * rng.random((ny, nx, nz))).astype(ft)
    vs = (vp / ft(1.8)).astype(ft)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(ft)
    meta = dict(dx=dx, freq=0.25, pmld=750.0, pmlr=1e-3, tau=tau)
    ierr, ir, jc, ref = O.ref_assemble(nx, ny, nz, dx, 0.25, vp, vs, rho, 750.0, 1e-3, tau=tau, prec=prec)
    assert ierr == 0
    dx_, om, pmld, pmlf, ipml = ref_params(meta, prec)
    assert om.imag != 0.0
    lam, mu, rho1 = lame_ref(vp, vs, rho, prec)
    rc, a = O.assemble(nx, ny, nz, dx_, om, lam, mu, rho1, pmld, pmlf, ipml, prec=prec)
    rel, zm = rel_err_entries(a, ref)
    assert rc == 0 and zm == 0 and rel <= (1e-12 if prec == 64 else 3e-5)
    # and it is not the real-frequency operator
    _, _, _, ref_real = O.ref_assemble(nx, ny, nz, dx, 0.25, vp, vs, rho, 750.0, 1e-3, prec=prec)
    assert rel_err_entries(ref_real, ref)[0] > 1e-3


# ---- physics invariants: pins that do not rely on the reference ----

def _interior_rows(nx, ny, nz, ipml):
    g = np.arange(nx * ny * nz)
    iz, ix, iy = g % nz + 1, (g // nz) % nx + 1, g // (nx * nz) + 1
    inter = (ix > ipml) & (ix <= nx - ipml) & (iy > ipml) & (iy <= ny - ipml) & \
            (iz > ipml) & (iz <= nz - ipml)
    return inter


def _matrix(nx, ny, nz, dx, omega, vp, vs, rho, pmld=500.0, fst=True):
    lam, mu, rho1 = lame_ref(vp, vs, rho, 64)
    ipml, pmlf = O.pml_params(dx, pmld, 1e-3, 64)
    rc, a = O.assemble(nx, ny, nz, dx, omega, lam, mu, rho1, pmld, pmlf, ipml, fst=fst, prec=64)
    assert rc == 0
    ir, jc = O.graph(nx, ny, nz)
    n = 3 * nx * ny * nz
    return sp.csr_matrix((a, jc - 1, ir - 1), shape=(n, n)), ipml


def test_rigid_translation_static():
    """At omega -> 0 the interior stiffness annihilates a rigid translation,
    also in a heterogeneous medium (small omega keeps the PML defined)."""
    nx, ny, nz = 11, 10, 12
    rng = np.random.default_rng(3)
    vp = 4000 + 2000 * rng.random((ny, nx, nz))
    vs = vp / (1.7 + 0.1 * rng.random((ny, nx, nz)))
    rho = 2000 + 800 * rng.random((ny, nx, nz))
    A, ipml = _matrix(nx, ny, nz, 250.0, 1e-9, vp, vs, rho)
    inter = _interior_rows(nx, ny, nz, ipml)
    scale = abs(A).max()
    for c in range(3):
        u = np.zeros(A.shape[0])
        u[c::3] = 1.0
        r = (A @ u).reshape(-1, 3)[inter]
        assert np.abs(r).max() <= 1e-12 * scale


def test_mass_row_sum_uniform():
    """Uniform medium: interior rows applied to a rigid translation give
    omega^2 rho (the lumping weights sum to one)."""
    nx, ny, nz = 11, 10, 12
    om = 2 * np.pi * 0.3
    vp = np.full((ny, nx, nz), 5000.0)
    A, ipml = _matrix(nx, ny, nz, 250.0, om, vp, vp / 1.8, np.full_like(vp, 2500.0))
    inter = _interior_rows(nx, ny, nz, ipml)
    for c in range(3):
        u = np.zeros(A.shape[0])
        u[c::3] = 1.0
        r = (A @ u).reshape(-1, 3)[inter]
        assert np.allclose(r[:, c], om * om * 2500.0, rtol=1e-10, atol=0)
        others = [d for d in range(3) if d != c]
        assert np.abs(r[:, others]).max() <= 1e-10 * om * om * 2500.0


def test_interior_symmetry_homogeneous():
    nx, ny, nz = 11, 10, 12
    vp = np.full((ny, nx, nz), 5000.0)
    A, ipml = _matrix(nx, ny, nz, 250.0, 2 * np.pi * 0.2, vp, vp / 1.8, np.full_like(vp, 2500.0))
    inter = np.where(_interior_rows(nx, ny, nz, ipml))[0]
    idx = (3 * inter[:, None] + np.arange(3)[None, :]).ravel()
    S = A[idx][:, idx]
    assert abs(S - S.T).max() <= 1e-12 * abs(S).max()


def test_dirichlet_rows_identity():
    nx, ny, nz = 7, 6, 8
    vp = np.full((ny, nx, nz), 5000.0)
    A, ipml = _matrix(nx, ny, nz, 250.0, 1.0, vp, vp / 1.8, np.full_like(vp, 2500.0))
    g = np.arange(nx * ny * nz)
    iz, ix, iy = g % nz + 1, (g // nz) % nx + 1, g // (nx * nz) + 1
    fixed = (iz == nz) | (ix == 1) | (ix == nx) | (iy == 1) | (iy == ny)
    rows = (3 * np.where(fixed)[0][:, None] + np.arange(3)[None, :]).ravel()
    B = A[rows]
    assert np.array_equal(B.toarray(), sp.identity(A.shape[0], format="csr")[rows].toarray())


def test_traction_free_rows_give_half_the_surface_traction():
    """TPFREE (tpfree.f:55-420): at a top node outside the horizontal PML of
    a homogeneous model the rows of a linear displacement field are half the
    traction (sigma_xz, sigma_yz, sigma_zz) it exerts on the surface, a
    rigid translation gives zero rows, and there is no mass term."""
    nx, ny, nz, dx = 9, 9, 6, 250.0
    vp, vs, rho = 6000.0, 3500.0, 2600.0
    mu = vs * vs * rho
    lam = vp * vp * rho - 2 * mu
    sh = (ny, nx, nz)
    L, M = O.lame(np.full(sh, vp), np.full(sh, vs), np.full(sh, rho), 64)
    ipml, pmlf = O.pml_params(dx, 2 * dx, 1e-3, 64)
    rows = {}
    # (omega = 0 makes r1 = 0/0 at every node, in the reference as here)
    for om in (2 * np.pi * 0.05, 2 * np.pi * 0.4):
        rc, a = O.assemble(nx, ny, nz, dx, om, L, M, np.full(nx * ny * nz, rho), 2 * dx, pmlf, ipml, fst=True,
                           usemin=False, prec=64)
        assert rc == 0
        ir, jc = O.graph(nx, ny, nz)
        rows[om] = sp.csr_matrix((a, jc - 1, ir - 1), shape=(3 * nx * ny * nz,) * 2)
    A = rows[2 * np.pi * 0.05]
    g = np.arange(nx * ny * nz)
    iz, ix, iy = g % nz, (g // nz) % nx, g // (nx * nz)
    x, y, z = ix * dx, iy * dx, iz * dx
    o = 0 * x
    c = int(np.nonzero((iz == 0) & (ix == 4) & (iy == 4))[0][0])
    # no frequency dependence at a top node away from the PML (no mass term)
    assert abs(rows[2 * np.pi * 0.4][3 * c:3 * c + 3] - A[3 * c:3 * c + 3]).max() == 0
    F = lambda a, b, cc: np.stack([a, b, cc], 1).ravel()  # noqa: E731
    # field -> traction (sigma_xz, sigma_yz, sigma_zz)
    cases = [(F(1 + o, o, o), (0, 0, 0)), (F(o, o, 1 + o), (0, 0, 0)), (F(y, o, o), (0, 0, 0)),
             (F(z, o, o), (mu, 0, 0)), (F(o, z, o), (0, mu, 0)), (F(o, o, z), (0, 0, lam + 2 * mu)),
             (F(x, o, o), (0, 0, lam)), (F(o, y, o), (0, 0, lam)), (F(o, o, x), (mu, 0, 0)),
             (F(o, o, y), (0, mu, 0))]
    for u, t in cases:
        r = (A @ u)[3 * c:3 * c + 3]
        assert np.allclose(r, 0.5 * np.asarray(t, float), rtol=1e-12, atol=1e-12 * (lam + 2 * mu)), (r, t)
