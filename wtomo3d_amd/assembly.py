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
