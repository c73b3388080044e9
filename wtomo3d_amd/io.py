"""Project ingestion: the reference's ini parameters and HDF5 model files.

A project ``name`` is the pair ``name.ini`` (READINI_FORWARD,
readini.f90:34-222) and ``name_model.h5`` (MODELIO_READ_MODEL,
modelio.f90:1-92): float datasets /Model/Vp, /Model/Vs and /Model/Density
of nx*ny*nz values with x fastest, then y, then z (modelio.f90:98-113),
returned here as arrays (ny, nx, nz) with z fastest -- the layout of the
reference's MODEL_MODULE arrays (nz, nx, ny) and of the solver.

HDF5 is read and written through the HDF5 C library of this image (found by
``ctypes.util.find_library`` or under /opt/conda/lib); the ini file with the
standard-library parser, keys and defaults as READINI_FORWARD has them.
"""
import configparser
import ctypes
import ctypes.util
import math
import os

import numpy as np

__all__ = ["read_ini", "read_model_h5", "write_model_h5", "load_project", "ProjectError"]


class ProjectError(RuntimeError):
    pass


# ---------------------------------------------------------------- HDF5 ----

_H5 = None


def _h5():
    global _H5
    if _H5 is not None:
        return _H5
    cands = [ctypes.util.find_library("hdf5"), "/opt/conda/lib/libhdf5.so", "libhdf5.so"]
    lib = None
    for c in cands:
        if not c:
            continue
        try:
            lib = ctypes.CDLL(c)
            break
        except OSError:
            continue
    if lib is None:
        raise ProjectError("the HDF5 C library is not available")
    hid, herr, hsize = ctypes.c_int64, ctypes.c_int, ctypes.c_uint64
    lib.H5open.restype = herr
    lib.H5open()
    lib.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid]
    lib.H5Fopen.restype = hid
    lib.H5Fcreate.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid, hid]
    lib.H5Fcreate.restype = hid
    lib.H5Fclose.argtypes = [hid]
    lib.H5Gcreate2.argtypes = [hid, ctypes.c_char_p, hid, hid, hid]
    lib.H5Gcreate2.restype = hid
    lib.H5Gclose.argtypes = [hid]
    lib.H5Dopen2.argtypes = [hid, ctypes.c_char_p, hid]
    lib.H5Dopen2.restype = hid
    lib.H5Dcreate2.argtypes = [hid, ctypes.c_char_p, hid, hid, hid, hid, hid]
    lib.H5Dcreate2.restype = hid
    lib.H5Dget_space.argtypes = [hid]
    lib.H5Dget_space.restype = hid
    lib.H5Dread.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
    lib.H5Dread.restype = herr
    lib.H5Dwrite.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
    lib.H5Dwrite.restype = herr
    lib.H5Dclose.argtypes = [hid]
    lib.H5Screate_simple.argtypes = [ctypes.c_int, ctypes.POINTER(hsize), ctypes.POINTER(hsize)]
    lib.H5Screate_simple.restype = hid
    lib.H5Sget_simple_extent_npoints.argtypes = [hid]
    lib.H5Sget_simple_extent_npoints.restype = ctypes.c_int64
    lib.H5Sclose.argtypes = [hid]
    lib.H5Lexists.argtypes = [hid, ctypes.c_char_p, hid]
    lib.H5Lexists.restype = ctypes.c_int
    lib.native_float = hid.in_dll(lib, "H5T_NATIVE_FLOAT_g").value
    _H5 = lib
    return lib


def _read_float(lib, fid, name):
    did = lib.H5Dopen2(fid, name.encode(), 0)
    if did < 0:
        raise ProjectError("dataset %s not found" % name)
    try:
        sid = lib.H5Dget_space(did)
        n = lib.H5Sget_simple_extent_npoints(sid)
        lib.H5Sclose(sid)
        out = np.empty(n, np.float32)
        if lib.H5Dread(did, lib.native_float, 0, 0, 0, out.ctypes.data_as(ctypes.c_void_p)) < 0:
            raise ProjectError("cannot read %s" % name)
        return out
    finally:
        lib.H5Dclose(did)


def read_model_h5(path, nx, ny, nz):
    """(vp, vs, rho) as float32 arrays (ny, nx, nz), z fastest; raises when
    a dataset size is not nx*ny*nz (modelio.f90:46-52) or when /Model/Qp or
    /Model/Qs is present (the reference has no reader for them,
    modelio.f90:62-76)."""
    lib = _h5()
    if not os.path.exists(path):
        raise ProjectError("model file %s does not exist" % path)
    fid = lib.H5Fopen(path.encode(), 0, 0)
    if fid < 0:
        raise ProjectError("cannot open %s" % path)
    try:
        out = []
        for name in ("/Model/Vp", "/Model/Vs", "/Model/Density"):
            x = _read_float(lib, fid, name)
            if x.size != nx * ny * nz:
                raise ProjectError("%s holds %d values, the grid has %d" % (name, x.size, nx * ny * nz))
            # file index (iz-1) nx ny + (iy-1) nx + ix -> (iy, ix, iz)
            out.append(np.ascontiguousarray(x.reshape(nz, ny, nx).transpose(1, 2, 0)))
        for q in ("/Model/Qp", "/Model/Qs"):
            if lib.H5Lexists(fid, q.encode(), 0) > 0:
                raise ProjectError("%s present: attenuation models are passed to Solver.set_model_q" % q)
        return tuple(out)
    finally:
        lib.H5Fclose(fid)


def write_model_h5(path, vp, vs, rho):
    """Write (ny, nx, nz) z-fastest arrays as the reference's model file
    (1-D float datasets, x fastest, h5_cinter.c:143-170)."""
    lib = _h5()
    fid = lib.H5Fcreate(path.encode(), 2, 0, 0)  # H5F_ACC_TRUNC
    if fid < 0:
        raise ProjectError("cannot create %s" % path)
    try:
        gid = lib.H5Gcreate2(fid, b"/Model", 0, 0, 0)
        lib.H5Gclose(gid)
        for name, a in (("/Model/Vp", vp), ("/Model/Vs", vs), ("/Model/Density", rho)):
            flat = np.ascontiguousarray(np.asarray(a, np.float32).transpose(2, 0, 1)).ravel()
            dims = (ctypes.c_uint64 * 1)(flat.size)
            sid = lib.H5Screate_simple(1, dims, None)
            did = lib.H5Dcreate2(fid, name.encode(), lib.native_float, sid, 0, 0, 0)
            rc = lib.H5Dwrite(did, lib.native_float, 0, 0, 0, flat.ctypes.data_as(ctypes.c_void_p))
            lib.H5Dclose(did)
            lib.H5Sclose(sid)
            if rc < 0:
                raise ProjectError("cannot write %s" % name)
    finally:
        lib.H5Fclose(fid)


# ----------------------------------------------------------------- ini ----

_TRUE = {"y", "yes", "t", "true", "1"}
_FALSE = {"n", "no", "f", "false", "0"}


class _Ini:
    def __init__(self, path):
        cp = configparser.ConfigParser(interpolation=None, inline_comment_prefixes=(";", "#"))
        cp.optionxform = str.lower
        if not cp.read(path):
            raise ProjectError("cannot read %s" % path)
        self.cp = cp

    def raw(self, key):
        sec, opt = key.split(":")
        if self.cp.has_option(sec.lower(), opt.lower()):
            return self.cp.get(sec.lower(), opt.lower()).strip().strip('"')
        return None

    def int(self, key, default):
        v = self.raw(key)
        return default if v is None else int(float(v))

    def float(self, key, default):
        v = self.raw(key)
        return default if v is None else float(v)

    def bool(self, key, default):
        v = self.raw(key)
        if v is None:
            return default
        if v.lower() in _TRUE:
            return True
        if v.lower() in _FALSE:
            return False
        raise ProjectError("%s: not a boolean: %r" % (key, v))

    def wxyz(self, key):
        v = self.raw(key)
        if v is None:
            raise ProjectError("%s missing" % key)
        parts = v.replace(",", " ").split()
        if len(parts) != 4:
            raise ProjectError("%s: expected weight, x, y, z" % key)
        return tuple(float(p) for p in parts)


def read_ini(path):
    """The forward-modelling parameters of READINI_FORWARD (readini.f90:34-222)
    with its defaults and checks; raises ProjectError where it sets ierr.
    Station lines are 'weight, x, y, z'; positions are shifted by the model
    origin as the reference does."""
    ini = _Ini(path)
    p = {}
    nx, ny, nz = ini.int("model:nx", 0), ini.int("model:ny", 0), ini.int("model:nz", 0)
    if nx < 1 or ny < 1 or nz < 1:
        raise ProjectError("invalid nx, ny, nz: %d %d %d" % (nx, ny, nz))
    p.update(nx=nx, ny=ny, nz=nz)
    p["xorig"], p["yorig"], p["zorig"] = (ini.float("model:%s" % k, 0.0) for k in ("xorig", "yorig", "zorig"))
    dx = ini.float("model:dx", 0.0)
    dy, dz = ini.float("model:dy", dx), ini.float("model:dz", dx)
    if dx <= 0 or dy <= 0 or dz <= 0:
        raise ProjectError("invalid dx, dy, dz")
    if dx != dy or dx != dz:
        raise ProjectError("irregular grid spacing not permitted")
    p["dx"] = dx
    p["vpvs"] = ini.float("model:vpvs", math.sqrt(3.0))
    p["azm"] = ini.float("model:azm", 0.0)
    p["freqbase"] = ini.float("model:freqbase", 20.0)
    # freesurf(1..6): top, +x, bottom, -x, +y, -y
    p["freesurf"] = [ini.bool("model:fst", True), ini.bool("model:fsr", False), ini.bool("model:fsb", False),
                     ini.bool("model:fsl", False), ini.bool("model:fsm", False), ini.bool("model:fsp", False)]
    if not ini.bool("pml:pml", True):
        raise ProjectError("PMLs only")
    p["pmlr"] = ini.float("pml:pmlr", 0.001)
    p["pmld"] = ini.float("pml:pmld", 10.0 * dx)
    p["ipml"] = int(p["pmld"] / dx) + 1
    p["pmlf"] = 3.0 * math.log(1.0 / p["pmlr"]) / (2.0 * p["pmld"] ** 3)
    nom = ini.int("model:nom", 0)
    if nom < 1:
        raise ProjectError("no frequencies to model")
    freqs = []
    for k in range(1, nom + 1):
        f = ini.float("model:freq_%d" % k, 0.0)
        if f <= 0.0:
            raise ProjectError("error reading frequency %d" % k)
        freqs.append(f)
    p["freq"] = freqs
    nrgg = 9

    def stations(kind, count, prefix, origin):
        st = [ini.wxyz("receivers:%s_%d" % (prefix, i)) for i in range(1, count + 1)]
        w = np.array([s[0] for s in st])
        xyz = np.array([[s[1] + origin[0], s[2] + origin[1], s[3] + origin[2]] for s in st]).reshape(-1, 3)
        return w, xyz

    org = (p["xorig"], p["yorig"], p["zorig"])
    ng = ini.int("receivers:ng", 0)
    p["ng"] = ng
    gw, gxyz = stations("geo", ng, "geo", org) if ng > 0 else (np.zeros(0), np.zeros((0, 3)))
    p["igreg"] = ini.int("receivers:igreg", 0)
    p["gspread"] = ini.float("receivers:gspread", 0.5)
    p["usegwt"] = ini.bool("receivers:usegwt", False)
    if ng > 0 and ((p["gspread"] <= 0.5 and 2 * p["igreg"] + 1 > nrgg) or
                   (p["gspread"] > 0.5 and int(4.0 * p["gspread"]) + 1 > nrgg)):
        raise ProjectError("nrgg too small for geophones")
    p["gwght"] = gw if p["usegwt"] else np.ones(ng)
    p["geophones"] = gxyz
    nr = ini.int("receivers:nr", 0)
    p["nr"] = nr
    rw, rxyz = stations("rec", nr, "rec", org) if nr > 0 else (np.zeros(0), np.zeros((0, 3)))
    p["irreg"] = ini.int("receivers:irreg", p["igreg"])
    p["recin"] = ini.int("receivers:recin", 3)
    p["rspread"] = ini.float("receivers:rspread", p["gspread"])
    p["userwt"] = ini.bool("receivers:userwt", False)
    if nr > 0 and ((p["rspread"] <= 0.5 and 2 * p["irreg"] + 1 > nrgg) or
                   (p["rspread"] > 0.5 and int(4.0 * p["rspread"]) + 1 > nrgg)):
        raise ProjectError("nrgg too small for receivers")
    p["rwght"] = rw if p["userwt"] else np.ones(nr)
    p["hydrophones"] = rxyz
    ns = ini.int("sources:ns", 0)
    if ns < 1:
        raise ProjectError("no sources")
    p["ns"] = ns
    p["isreg"] = ini.int("sources:irreg", 0)
    p["srcin"] = ini.int("sources:srcin", 3)
    p["sspread"] = ini.float("sources:sspread", 0.5)
    if (p["sspread"] <= 0.5 and 2 * p["isreg"] + 1 > nrgg) or \
            (p["sspread"] > 0.5 and int(4.0 * p["sspread"]) + 1 > nrgg):
        raise ProjectError("nrgg too small for sources")
    if ini.int("sources:nsg", 1) != 1:
        raise ProjectError("nsg = 1 only")
    p["tau"] = ini.float("optional:tau", 999.999)
    p["usemin"] = ini.bool("optional:usemin", True)
    return p


def load_project(name):
    """(parameters, vp, vs, rho) of the project ``name`` (``name.ini`` and
    ``name_model.h5``)."""
    p = read_ini(name + ".ini")
    vp, vs, rho = read_model_h5(name + "_model.h5", p["nx"], p["ny"], p["nz"])
    return p, vp, vs, rho
