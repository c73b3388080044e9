"""Attenuation pre-pass oracle (oracle/model_oracle.py) against the
reference's DISPERSION and closed forms."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import model_oracle as MO  # noqa: E402

needs_ref = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built")


def attenuating_model(seed=4):
    rng = np.random.default_rng(seed)
    shape = (5, 6, 7)
    vp = 4000 + 3000 * rng.random(shape)
    vs = vp / (1.6 + 0.3 * rng.random(shape))
    qpi = rng.uniform(0.002, 0.05, shape)
    qsi = rng.uniform(0.002, 0.08, shape)
    qsi[0, 0, 0] = 25.0  # 1 + f q / pi < 0.1: the reference floors the factor
    rho = 2000 + 1000 * rng.random(shape)
    return vp, vs, qpi, qsi, rho


@needs_ref
@pytest.mark.parametrize("freqbase,omega", [(1.0, 2 * np.pi * 0.3), (0.5, 2 * np.pi * 2.0), (1.0, 0.0),
                                            (0.0, 2.0)])
def test_dispersion_matches_reference(freqbase, omega):
    vp, vs, qpi, qsi, rho = attenuating_model()
    lam, mu = MO.dispersion(vp, vs, qpi, qsi, rho, freqbase, omega)
    rl, rm = O.ref_dispersion(vp, vs, qpi, qsi, rho, freqbase, omega, prec=64)
    np.testing.assert_allclose(lam, rl.ravel().reshape(lam.shape), rtol=1e-14, atol=0)
    np.testing.assert_allclose(mu, rm.ravel().reshape(mu.shape), rtol=1e-14, atol=0)
    rl, rm = O.ref_dispersion(vp, vs, qpi, qsi, rho, freqbase, omega, prec=32)
    np.testing.assert_allclose(lam, rl, rtol=2e-6, atol=0)
    np.testing.assert_allclose(mu, rm, rtol=2e-6, atol=0)


def test_dispersion_closed_forms():
    vp, vs, qpi, qsi, rho = attenuating_model()
    # at the base frequency only the imaginary part of the speed changes
    lam, mu = MO.dispersion(vp, vs, qpi, qsi, rho, 1.0, 2 * np.pi)
    np.testing.assert_allclose(mu, (vs * (1 - 0.5j * qsi)) ** 2 * rho, rtol=1e-15)
    # lossless medium: the moduli of the base speeds at every frequency
    z = np.zeros_like(vp)
    lam, mu = MO.dispersion(vp, vs, z, z, rho, 1.0, 2 * np.pi * 0.2)
    np.testing.assert_allclose(mu, vs * vs * rho, rtol=1e-15)
    np.testing.assert_allclose(lam, vp * vp * rho - 2 * vs * vs * rho, rtol=1e-15)
    # the dispersion factor is floored at 0.1
    lam, mu = MO.dispersion(vp, vs, qpi, qsi, rho, 1.0, 2 * np.pi * 0.01)
    assert mu[0, 0, 0] == pytest.approx((0.1 * vs[0, 0, 0] * (1 - 12.5j)) ** 2 * rho[0, 0, 0], rel=1e-14)
