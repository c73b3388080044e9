"""Shared construction of small forward-solve test problems (oracle side)."""
import numpy as np

import oracle_lib as O


def layered_model(nx, ny, nz, dx, seed, perturb=0.0):
    """Layered Vp (increasing with depth), Vs = Vp/1.73, Gardner density,
    optional multiplicative random perturbation; arrays (ny, nx, nz)."""
    rng = np.random.default_rng(seed)
    z = np.arange(nz) * dx
    tops = np.sort(rng.uniform(0, nz * dx, 5))
    vps = np.sort(rng.uniform(4000, 8000, 6))
    vp1 = vps[np.searchsorted(tops, z, side="right")]
    vp = np.broadcast_to(vp1, (ny, nx, nz)).copy()
    vs = vp / 1.73
    if perturb:
        d = perturb * rng.standard_normal((ny, nx, nz))
        vp = vp * (1 + d)
        vs = vs * (1 + d)
        rho = 310.0 * (vp1[None, None, :] ** 0.25) * (1 + 0.3 * d)
    else:
        rho = np.broadcast_to(310.0 * vp1 ** 0.25, (ny, nx, nz)).copy()
    return vp, vs, rho


def problem(nx, ny, nz, dx=250.0, freq=0.3, seed=1, perturb=0.05, pmld=None):
    """(params tuple, lam, mu, rho, rowptr, colidx, vals) in complex128."""
    vp, vs, rho = layered_model(nx, ny, nz, dx, seed, perturb)
    lam, mu = O.lame(vp, vs, rho, 64)
    if pmld is None:
        pmld = 2 * dx
    ipml, pmlf = O.pml_params(dx, pmld, 1e-3, 64)
    om = 2 * np.pi * freq
    rc, vals = O.assemble(nx, ny, nz, dx, om, lam, mu, rho.ravel(), pmld, pmlf, ipml, prec=64)
    assert rc == 0
    ir, jc = O.graph(nx, ny, nz)
    return dict(dx=dx, omega=om, pmld=pmld, pmlf=pmlf, ipml=ipml), lam, mu, rho.ravel(), ir, jc, vals


def rhs_block(A, nrhs, seed):
    """B = A X_true for a random X_true: right-hand sides scaled like the
    operator, as a physical source term is."""
    rng = np.random.default_rng(seed)
    n = A.shape[0]
    X = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    return np.asarray(A @ X), X
