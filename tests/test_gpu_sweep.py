"""Frequency sweep of a layered model with Thomson-Haskell plane-wave
sources (the reference's frequency loop, elwave.f90:147-186, with its empty
solve phase completed).  The right-hand sides are A u for the plane-wave
fields u, so at every frequency the refined solution must reproduce u, and
agree with the complex128 factor's solution."""
import numpy as np
import pytest

from wtomo3d_amd import sweep

pytestmark = pytest.mark.gpu


def test_layered_sweep_recovers_the_plane_waves():
    w = sweep.layered_workload(nx=14, ny=12, nz=10, dx=250.0, seed=1002, freqs=[0.12, 0.3, 0.45], nsrc=4,
                               pml_nodes=3)
    recs = sweep.sweep(w, prec=32, tol=1e-10, max_iter=30, compare_fp64=True)
    assert [r["freq"] for r in recs] == [0.12, 0.3, 0.45]
    for r in recs:
        assert r["rel_residual"] <= 1e-10 and r["iterations"] >= 2
        assert r["rel_wavefield_err"] <= 1e-7, r
        assert r["fp64_rel_residual"] <= 1e-10  # one complex128 solve, no refinement
        assert r["vs_fp64_rel_diff"] <= 1e-7, r



def test_frequency_batched_sweep_matches_the_sequential_one():
    """All frequencies factored together (one block-diagonal system, one or
    several groups) give the same plane waves as the frequency loop."""
    w = sweep.layered_workload(nx=14, ny=12, nz=10, dx=250.0, seed=1002, freqs=[0.12, 0.3, 0.45], nsrc=4,
                               pml_nodes=3)
    seq = sweep.sweep(w, prec=32, tol=1e-10, max_iter=30)
    for gb in (120e9, 1.0):  # one group; one frequency per group
        bat = sweep.sweep_batched(w, prec=32, tol=1e-10, max_iter=30, max_group_bytes=gb)
        assert bat["groups"] == (1 if gb > 1 else 3)
        assert [r["freq"] for r in bat["per_frequency"]] == [0.12, 0.3, 0.45]
        for r, q in zip(bat["per_frequency"], seq):
            assert r["rel_residual"] <= 1e-10 and r["iterations"] >= 2, r
            assert r["rel_wavefield_err"] <= 1e-7, r
            assert abs(r["rel_wavefield_err"] - q["rel_wavefield_err"]) <= 1e-7
