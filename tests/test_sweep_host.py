"""The layered multi-frequency workload generator (host only)."""
import numpy as np

from wtomo3d_amd import sweep


def test_workload_is_seeded_and_shaped():
    a = sweep.layered_workload(nx=8, ny=8, nz=6, nsrc=5)
    b = sweep.layered_workload(nx=8, ny=8, nz=6, nsrc=5)
    assert a["freqs"] == [0.1, 0.15, 0.2, 0.25, 0.3, 0.35, 0.4, 0.45]
    assert len(a["sources"]) == 8 and all(len(s["ain"]) == 5 for s in a["sources"])
    for sa, sb in zip(a["sources"], b["sources"]):
        assert np.array_equal(sa["ain"], sb["ain"]) and np.array_equal(sa["phi"], sb["phi"])
        assert np.all((sa["ain"] >= np.deg2rad(10)) & (sa["ain"] <= np.deg2rad(35)))
