"""Plane-wave right-hand sides of a layered half-space, on the device.

The reference computes 1-D Thomson-Haskell depth profiles per plane-wave
source with HASKGRN (haskgrn.f:31-672) and the scattered-field source
(A_bg - A) u_bg with HETFLDSRC (hetfldsrc.f90:17-1039); the expansion of the
profiles to 3-D fields (planewvexp.f90) is missing from the reference and
is defined here as

    u_j = (U_j(z) cos phi_j, U_j(z) sin phi_j, W_j(z))
          * exp(-i kh_j (x cos phi_j + y sin phi_j)),   kh_j = omega / cv_j

(phi_j the azimuth of horizontal propagation from +x towards +y, z down).

    prof = profiles(layers, ain, is_s, omega, nz, dz)      # HASKGRN, ns sources
    ubg = background_fields(nx, ny, nz, dx, prof, phi)     # device block (n, ns)
    sfld = solver.scattered_source(lam_bg, mu_bg, rho_bg, ubg, srcfn)
"""
import ctypes

import numpy as np

from ._lib import lib, check

_vp = ctypes.c_void_p
WT_PREC_F32, WT_PREC_F64 = 32, 64


class DeviceBuffer:
    """A device allocation of the library (freed with the object)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(lib().wt_dev_alloc(max(self.nbytes, 1), ctypes.byref(self.ptr)), "wt_dev_alloc")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(lib().wt_h2d(b.ptr, a.ctypes.data_as(_vp), a.nbytes), "wt_h2d")
        return b

    def to_host(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        check(lib().wt_d2h(out.ctypes.data_as(_vp), self.ptr, out.nbytes), "wt_d2h")
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().wt_dev_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Profiles:
    """Device-resident depth profiles of ns sources: U, W [ns][nz] and kh[ns]."""

    def __init__(self, U, W, kh, ns, nz, prec, status):
        self.U, self.W, self.kh = U, W, kh
        self.ns, self.nz, self.prec = ns, nz, prec
        self.status = status

    def host(self):
        ct = np.complex64 if self.prec == WT_PREC_F32 else np.complex128
        return (self.U.to_host(ct, (self.ns, self.nz)), self.W.to_host(ct, (self.ns, self.nz)),
                self.kh.to_host(np.float64, (self.ns,)))


def profiles(layers, ain, is_s, omega, nz, dz, zorig=0.0, prec=WT_PREC_F64, stream=None, out=None):
    """HASKGRN for every source on the device.

    layers: dict of per-layer sequences ``vp, vs, rho, top, vpd, vsd`` (the
    last layer the half-space); ain: incidence angles (radians); is_s: 1 for
    an incident S wave, 0 for P.  ``stream``: the HIP stream to run on (a
    solver's, ``Solver.stream()``); ``out``: (U, W, kh) DeviceBuffers to
    reuse.  Raises on a reference error (a depth above the first layer top)."""
    ft = np.float32 if prec == WT_PREC_F32 else np.float64
    ct = np.complex64 if prec == WT_PREC_F32 else np.complex128
    arrs = [np.ascontiguousarray(layers[k], ft) for k in ("vp", "vs", "rho", "top", "vpd", "vsd")]
    ain = np.ascontiguousarray(np.atleast_1d(ain), ft)
    ns = ain.size
    iss = np.ascontiguousarray(np.broadcast_to(np.asarray(is_s, np.int32), (ns,)))
    nb = ns * nz * np.dtype(ct).itemsize
    if out is not None and out[0].nbytes >= nb and out[1].nbytes >= nb and out[2].nbytes >= ns * 8:
        U, W, kh = out
    else:
        U, W, kh = DeviceBuffer(nb), DeviceBuffer(nb), DeviceBuffer(ns * 8)
    status = np.zeros(ns, np.int32)
    check(lib().wt_haskgrn(prec, arrs[0].size, *[a.ctypes.data_as(_vp) for a in arrs], ns,
                           ain.ctypes.data_as(_vp), iss.ctypes.data_as(_vp), float(omega), nz, float(dz),
                           float(zorig), U.ptr, W.ptr, kh.ptr, status.ctypes.data_as(_vp), stream), "wt_haskgrn")
    if status.any():
        raise RuntimeError("haskgrn: depth above the first layer top for sources %s" % np.nonzero(status)[0])
    return Profiles(U, W, kh, ns, nz, prec, status)


def background_fields(nx, ny, nz, dx, prof, phi, stream=None, out=None):
    """3-D background fields of the profiled plane waves: a device block,
    row-major (n, ns) complex128 in the unknown order (into ``out``, a
    DeviceBuffer, when given)."""
    assert prof.nz == nz
    phi = np.ascontiguousarray(np.broadcast_to(np.asarray(phi, np.float64), (prof.ns,)))
    nb = 3 * nx * ny * nz * prof.ns * 16
    X = out if out is not None and out.nbytes >= nb else DeviceBuffer(nb)
    check(lib().wt_planewave_layered(nx, ny, nz, float(dx), prof.prec, prof.ns, prof.U.ptr, prof.W.ptr,
                                     prof.kh.ptr, phi.ctypes.data_as(_vp), X.ptr, stream), "wt_planewave_layered")
    return X


def total_field(uh, ubg, count, stream=None):
    """uh += ubg over ``count`` complex128 values (DeviceBuffers): the total
    field of MAKEUH (makeuh.f:41-59)."""
    check(lib().wt_total_field(uh.ptr, ubg.ptr, int(count), stream), "wt_total_field")


class Receivers:
    """Receivers and hydrophones of one grid as a device extraction matrix
    (MAKEUH's receiver and hydrophone loops, makeuh.f:60-168).

    rec / hyd: dicts x, y, z (metres from the first node), spread, ireg
    (rec may carry ``fac``, complex responses); hydrophones need the
    model's lam, mu (complex, z fastest) and omega.  ``extract(field, nrhs)``
    returns (uest, vest, west, utest) arrays of shape (count, nrhs)."""

    def __init__(self, nx, ny, nz, dx, rec=None, hyd=None, lam=None, mu=None, omega=0.0,
                 freesurf=(1, 0, 0, 0, 0, 0)):
        rec = rec or dict(x=[], y=[], z=[], spread=0.5, ireg=0)
        hyd = hyd or dict(x=[], y=[], z=[], spread=0.5, ireg=0)
        A = lambda v: np.ascontiguousarray(v, np.float64)  # noqa: E731
        self._keep = [A(rec[k]) for k in ("x", "y", "z")] + [A(hyd[k]) for k in ("x", "y", "z")]
        self.ng, self.nr = len(rec["x"]), len(hyd["x"])
        fac = None
        if rec.get("fac") is not None:
            fac = np.ascontiguousarray(rec["fac"], np.complex128)
            self._keep.append(fac)
        fs = np.ascontiguousarray(freesurf, np.int32)
        lm = [np.ascontiguousarray(a, np.complex128).ravel() if a is not None else None for a in (lam, mu)]
        self._keep += [fs] + [a for a in lm if a is not None]
        om = complex(omega)
        h = ctypes.c_void_p()
        P = lambda a: None if a is None else a.ctypes.data_as(_vp)  # noqa: E731
        k = self._keep
        check(lib().wt_receivers_create(nx, ny, nz, float(dx), P(fs), self.ng, P(k[0]), P(k[1]), P(k[2]),
                                        float(rec["spread"]), int(rec["ireg"]), P(fac), self.nr, P(k[3]), P(k[4]),
                                        P(k[5]), float(hyd["spread"]), int(hyd["ireg"]), P(lm[0]), P(lm[1]),
                                        om.real, om.imag, ctypes.byref(h)), "wt_receivers_create")
        self._h = h
        self.rows = lib().wt_receivers_rows(h)
        self.n = 3 * nx * ny * nz

    def extract(self, field, nrhs=None, stream=None):
        host = not isinstance(field, DeviceBuffer)
        if host:
            field = np.ascontiguousarray(np.asarray(field).reshape(self.n, -1), np.complex128)
            nrhs = field.shape[1]
            dev = DeviceBuffer.from_host(field)
        else:
            nrhs = int(nrhs) if nrhs is not None else field.nbytes // (self.n * 16)
            dev = field
        out = DeviceBuffer(self.rows * nrhs * 16)
        check(lib().wt_receivers_extract(self._h, nrhs, dev.ptr, out.ptr, stream), "wt_receivers_extract")
        E = out.to_host(np.complex128, (self.rows, nrhs))
        g = E[:3 * self.ng].reshape(self.ng, 3, nrhs)
        return g[:, 0], g[:, 1], g[:, 2], E[3 * self.ng:]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().wt_receivers_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
