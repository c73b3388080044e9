"""GPU attenuation pre-pass (wt_dispersion, wt_solver_set_model_q) against
the oracle: the same double operations, 1e-14 relative."""
import os
import sys

import numpy as np
import pytest

from test_model_oracle import attenuating_model
from wtomo3d_amd import media
from wtomo3d_amd._lib import grid_params
from wtomo3d_amd.solver import Solver

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import model_oracle as MO  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("freqbase,omega", [(1.0, 2 * np.pi * 0.3), (0.5, 2 * np.pi * 2.0), (1.0, 0.0)])
def test_dispersion_matches_oracle(freqbase, omega):
    vp, vs, qpi, qsi, rho = attenuating_model()
    lam, mu, _ = media.attenuating(vp, vs, qpi, qsi, rho, freqbase, omega)
    rl, rm = MO.dispersion(vp, vs, qpi, qsi, rho, freqbase, omega)
    np.testing.assert_allclose(lam, rl.ravel(), rtol=1e-14, atol=0)
    np.testing.assert_allclose(mu, rm.ravel(), rtol=1e-14, atol=0)


def test_solver_model_with_attenuation():
    """set_model_q then assemble gives the operator of set_model with the
    oracle's complex moduli."""
    vp, vs, qpi, qsi, rho = attenuating_model(6)
    ny, nx, nz = vp.shape
    om = 2 * np.pi * 0.35
    p = grid_params(nx, ny, nz, 300.0, om, pmld=1200.0)
    a = Solver(nx, ny, nz)
    a.set_model_q(vp, vs, qpi, qsi, rho, 1.0, om)
    a.assemble(p)
    b = Solver(nx, ny, nz)
    rl, rm = MO.dispersion(vp, vs, qpi, qsi, rho, 1.0, om)
    b.set_model(rl.ravel(), rm.ravel(), rho.ravel())
    b.assemble(p)
    va, vb = a.values(), b.values()
    assert np.abs(va - vb).max() <= 1e-13 * np.abs(vb).max()
    a.close()
    b.close()
