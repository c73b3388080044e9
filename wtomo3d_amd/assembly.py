"""CSR pattern and stencil assembly (host side of ASMBLE, asmble.f90:1).

Mirrors the reference's call sequence on the host: SREL3D/SPARSE_ADJ2CRS
(elwave.f90:92-124) -> ``graph``; ASMBLE (elwave.f90:171) -> ``assemble``.
"""
import ctypes
import numpy as np

from ._lib import lib, check

_vp = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(_vp)


def graph_nnz(nx, ny, nz):
    return int(lib().wt_graph_nnz(nx, ny, nz))


def graph(nx, ny, nz, base=1):
    """CSR (rowptr, colidx) of the 27-point x 3-component graph, built on the GPU."""
    n = 3 * nx * ny * nz
    rowptr = np.zeros(n + 1, np.int32)
    colidx = np.zeros(graph_nnz(nx, ny, nz), np.int32)
    check(lib().wt_graph(nx, ny, nz, base, _p(rowptr), _p(colidx)), "wt_graph")
    return rowptr, colidx


def lame_from_velocity(vp, vs, rho, dtype=np.complex128):
    """lambda, mu as MODELIO_READ_MODEL forms them (modelio.f90:78-92)."""
    rt = np.float32 if dtype == np.complex64 else np.float64
    vp = np.asarray(vp, rt)
    vs = np.asarray(vs, rt)
    rho = np.asarray(rho, rt)
    mu = (vs * vs) * rho
    lam = (vp * vp) * rho - rt(2) * mu
    return lam.astype(dtype), mu.astype(dtype)


def assemble(params, lam, mu, rho, prec=64):
    """CSR values on the GPU for one frequency; returns a complex array[nnz]."""
    ct = np.complex64 if prec == 32 else np.complex128
    rt = np.float32 if prec == 32 else np.float64
    lam = np.ascontiguousarray(lam, ct).ravel()
    mu = np.ascontiguousarray(mu, ct).ravel()
    rho = np.ascontiguousarray(rho, rt).ravel()
    nodes = params.nx * params.ny * params.nz
    if lam.size != nodes or mu.size != nodes or rho.size != nodes:
        raise ValueError("model arrays must have nx*ny*nz entries")
    vals = np.zeros(graph_nnz(params.nx, params.ny, params.nz), ct)
    check(lib().wt_assemble(ctypes.byref(params), prec, _p(lam), _p(mu), _p(rho), _p(vals)),
          "wt_assemble")
    return vals


def _main():
    """Assembly rate on the device-resident operator of a heterogeneous model:
    ``python -m wtomo3d_amd.assembly [--nx 96 --ny 96 --nz 64 --reps 20]``.
    Prints one JSON line: ms per launch from HIP events on the solver stream
    and the algorithmic bytes (complex128 values written, model read once)."""
    import argparse
    import json

    from . import media
    from ._lib import grid_params
    from .solver import Solver

    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=96)
    ap.add_argument("--ny", type=int, default=96)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--dx", type=float, default=200.0)
    ap.add_argument("--freq", type=float, default=0.3)
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args()
    vp, vs, rho = media.heterogeneous(a.nx, a.ny, a.nz, a.dx, seed=1003)
    lam, mu, rho1 = media.lame(vp, vs, rho)
    s = Solver(a.nx, a.ny, a.nz)
    s.set_model(lam, mu, rho1)
    g = grid_params(a.nx, a.ny, a.nz, a.dx, 2 * np.pi * a.freq, pmld=10 * a.dx)
    s.assemble(g)  # warm-up
    s.profile(True)
    for _ in range(a.reps):
        s.assemble(g)
    pr = s.profile_read()["assemble"]
    ms = pr["ms"] / max(pr["launches"], 1)
    gbs = pr["bytes"] / (pr["ms"] * 1e-3) / 1e9 if pr["ms"] > 0 else 0.0
    print(json.dumps({"workload": "FP64 assembly %dx%dx%d dx=%gm %g Hz" % (a.nx, a.ny, a.nz, a.dx, a.freq),
                      "nnz": graph_nnz(a.nx, a.ny, a.nz), "launches": pr["launches"], "ms_per_launch": ms,
                      "bytes_per_launch": pr["bytes"] / max(pr["launches"], 1), "gbs": gbs,
                      "frac_hbm": gbs / 8000.0}))
    s.close()


if __name__ == "__main__":
    _main()
