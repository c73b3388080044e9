"""GPU receiver extraction (wt_receivers_*) against the reference's MAKEUH
fixture and the oracle.  The device applies the same weights as a sparse
matrix in FP64: 1e-12 of the largest output."""
import numpy as np
import pytest

from test_rcv_oracle import golden_case
from wtomo3d_amd import rhs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", range(5))
def test_receivers_match_reference_fixture(k):
    g, (nx, ny, nz, dx, freq), (u, v, w), lam, mu, sets = golden_case()
    tag, rec, hyd = sets[k]
    R = rhs.Receivers(nx, ny, nz, dx, rec, hyd, lam, mu, 2 * np.pi * freq)
    field = np.stack([u.ravel(), v.ravel(), w.ravel()], axis=1).ravel()
    # three columns: the field, twice the field, the field times i
    X = np.stack([field, 2 * field, 1j * field], axis=1)
    ue, ve, we, pe = R.extract(X)
    out = np.concatenate([ue, ve, we, pe])
    ref = g[tag + "__out64"]
    for j, f in enumerate((1, 2, 1j)):
        assert np.abs(out[:, j] - f * ref).max() <= 1e-12 * np.abs(ref).max()
    R.close()


def test_receivers_reject_stations_off_the_grid():
    with pytest.raises(RuntimeError):
        rhs.Receivers(6, 6, 6, 100.0, dict(x=[900.0], y=[10.0], z=[10.0], spread=0.5, ireg=0))
    with pytest.raises(RuntimeError):
        rhs.Receivers(6, 6, 6, 100.0, dict(x=[100.0], y=[10.0], z=[10.0], spread=0.5, ireg=5))
