"""ctypes bindings for the CPU checkers (test infrastructure only).

- ``oracle/liborasm.so``: the plain-C restatement (oracle/asm_oracle.c).
- ``oracle/_ref/libwtref{32,64}.so``: the reference's own Fortran assembly
  sources compiled by oracle/Makefile (present only where they were built).
"""
import ctypes
import os
import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liborasm.so")
REF_SO = {32: os.path.join(ROOT, "oracle", "_ref", "libwtref32.so"),
          64: os.path.join(ROOT, "oracle", "_ref", "libwtref64.so")}

_c_int_p = ctypes.POINTER(ctypes.c_int)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


_ORA = None


def oracle():
    global _ORA
    if _ORA is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liborasm.so is not built (make -C oracle)")
        L = ctypes.CDLL(ORACLE_SO)
        L.ora_graph_size.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_longlong)] * 2
        L.ora_graph.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 2
        for sfx, R in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            f = getattr(L, "ora_assemble_" + sfx)
            f.argtypes = [ctypes.c_int] * 3 + [R] * 3 + [ctypes.c_int] * 2 + [R] * 2 + \
                [ctypes.c_int] + [ctypes.c_void_p] * 4
            f.restype = ctypes.c_int
            g = getattr(L, "ora_lame_" + sfx)
            g.argtypes = [ctypes.c_long] + [ctypes.c_void_p] * 5
            h = getattr(L, "ora_pml_params_" + sfx)
            h.argtypes = [R] * 3 + [_c_int_p, ctypes.POINTER(R)]
        _ORA = L
    return _ORA


def graph_size(nx, ny, nz):
    n = ctypes.c_longlong()
    nnz = ctypes.c_longlong()
    oracle().ora_graph_size(nx, ny, nz, ctypes.byref(n), ctypes.byref(nnz))
    return n.value, nnz.value


def graph(nx, ny, nz):
    n, nnz = graph_size(nx, ny, nz)
    ir = np.zeros(n + 1, np.int32)
    jc = np.zeros(nnz, np.int32)
    rc = oracle().ora_graph(nx, ny, nz, _ptr(ir), _ptr(jc))
    assert rc == 0
    return ir, jc


def pml_params(dx, pmld, pmlr, prec=64):
    R = ctypes.c_float if prec == 32 else ctypes.c_double
    ip = ctypes.c_int()
    pf = R()
    getattr(oracle(), "ora_pml_params_f%d" % prec)(dx, pmld, pmlr, ctypes.byref(ip), ctypes.byref(pf))
    return ip.value, pf.value


def lame(vp, vs, rho, prec=64):
    """(lambda, mu) as complex arrays from (nz,nx,ny)-ordered vp, vs, rho."""
    ft = np.float32 if prec == 32 else np.float64
    ct = np.complex64 if prec == 32 else np.complex128
    vp = np.ascontiguousarray(vp, ft).ravel()
    vs = np.ascontiguousarray(vs, ft).ravel()
    rho = np.ascontiguousarray(rho, ft).ravel()
    lam = np.zeros(vp.size, ct)
    mu = np.zeros(vp.size, ct)
    getattr(oracle(), "ora_lame_f%d" % prec)(vp.size, _ptr(vp), _ptr(vs), _ptr(rho), _ptr(lam), _ptr(mu))
    return lam, mu


def assemble(nx, ny, nz, dx, omega, lam, mu, rho, pmld, pmlf, ipml,
             fst=True, usemin=True, prec=64):
    ft = np.float32 if prec == 32 else np.float64
    ct = np.complex64 if prec == 32 else np.complex128
    _, nnz = graph_size(nx, ny, nz)
    a = np.zeros(nnz, ct)
    lam = np.ascontiguousarray(lam, ct)
    mu = np.ascontiguousarray(mu, ct)
    rho = np.ascontiguousarray(rho, ft)
    om = complex(omega)
    rc = getattr(oracle(), "ora_assemble_f%d" % prec)(
        nx, ny, nz, dx, om.real, om.imag, int(fst), int(usemin), pmld, pmlf, ipml,
        _ptr(lam), _ptr(mu), _ptr(rho), _ptr(a))
    return rc, a


def have_ref():
    return all(os.path.exists(p) for p in REF_SO.values())


_REF = {}


def ref(prec):
    if prec not in _REF:
        L = ctypes.CDLL(REF_SO[prec])
        _REF[prec] = L
    return _REF[prec]


def ref_graph(nx, ny, nz):
    L = ref(32)
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    nv = ctypes.c_int()
    nzr = ctypes.c_int()
    L.wtref_graph_size(I(nx), I(ny), I(nz), ctypes.byref(nv), ctypes.byref(nzr))
    ir = np.zeros(nv.value + 1, np.int32)
    jc = np.zeros(nzr.value, np.int32)
    L.wtref_graph(I(nx), I(ny), I(nz), _ptr(ir), _ptr(jc))
    return ir, jc


def ref_assemble(nx, ny, nz, dx, freq, vp, vs, rho, pmld, pmlr, fst=True, usemin=True,
                 tau=999.999, prec=64):
    """Reference ASMBLE on (nz,nx,ny)-ordered velocity/density arrays."""
    L = ref(prec)
    R = ctypes.c_float if prec == 32 else ctypes.c_double
    ft = np.float32 if prec == 32 else np.float64
    ct = np.complex64 if prec == 32 else np.complex128
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    F = lambda v: ctypes.byref(R(v))
    ir, jc = ref_graph(nx, ny, nz)
    a = np.zeros(jc.size, ct)
    ierr = ctypes.c_int(-1)
    vp = np.ascontiguousarray(vp, ft)
    vs = np.ascontiguousarray(vs, ft)
    rho = np.ascontiguousarray(rho, ft)
    L.wtref_assemble(I(nx), I(ny), I(nz), F(dx), F(freq), F(tau), I(int(fst)), I(int(usemin)),
                     F(pmld), F(pmlr), _ptr(vp), _ptr(vs), _ptr(rho), I(jc.size),
                     _ptr(ir), _ptr(jc), _ptr(a), ctypes.byref(ierr))
    return ierr.value, ir, jc, a


def _hk_arrays(layers, ft):
    """layers: dict of per-layer sequences vp, vs, rho, top, vpd, vsd."""
    return [np.ascontiguousarray(layers[k], ft) for k in ("vp", "vs", "rho", "top", "vpd", "vsd")]


def haskgrn(layers, ain, is_s, omega, nz, dz, zorig, prec=32):
    """Oracle HASKGRN (oracle/hask_oracle.c) -> (rc, ugrn, wgrn, pic)."""
    R = ctypes.c_float if prec == 32 else ctypes.c_double
    ft = np.float32 if prec == 32 else np.float64
    ct = np.complex64 if prec == 32 else np.complex128
    f = getattr(oracle(), "ora_haskgrn_f%d" % prec)
    f.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 6 + [R, ctypes.c_int, R, ctypes.c_int, R, R] + \
        [ctypes.c_void_p] * 2 + [ctypes.POINTER(ctypes.c_double)]
    f.restype = ctypes.c_int
    arrs = _hk_arrays(layers, ft)
    u = np.zeros(nz, ct)
    w = np.zeros(nz, ct)
    pic = ctypes.c_double()
    rc = f(len(arrs[0]), *[_ptr(a) for a in arrs], ain, int(is_s), omega, nz, dz, zorig, _ptr(u), _ptr(w),
           ctypes.byref(pic))
    return rc, u, w, pic.value


def ref_haskgrn(layers, ain, is_s, omega, nz, dz, zorig, prec=32):
    """Reference HASKGRN (haskgrn.f:31) through oracle/_ref -> (ierr, ugrn, wgrn)."""
    L = ref(prec)
    R = ctypes.c_float if prec == 32 else ctypes.c_double
    ft = np.float32 if prec == 32 else np.float64
    ct = np.complex64 if prec == 32 else np.complex128
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    F = lambda v: ctypes.byref(R(v))
    arrs = _hk_arrays(layers, ft)
    u = np.zeros(nz, ct)
    w = np.zeros(nz, ct)
    ierr = ctypes.c_int(-1)
    L.wtref_haskgrn(I(len(arrs[0])), *[_ptr(a) for a in arrs], F(ain), I(int(is_s)), F(omega), I(nz), F(dz),
                    F(zorig), _ptr(u), _ptr(w), ctypes.byref(ierr))
    return ierr.value, u, w


def ref_hetfldsrc(nx, ny, nz, dx, freq, vp, vs, rho, vpb, vsb, rhob, ub, vb, wb, srcfn, pmld, pmlr,
                  fst=True, usemin=True, tau=999.999, prec=64):
    """Reference HETFLDSRC (hetfldsrc.f90:17) -> (ierr, sfld[3 n]); grids (nz,nx,ny)."""
    L = ref(prec)
    R = ctypes.c_float if prec == 32 else ctypes.c_double
    ft = np.float32 if prec == 32 else np.float64
    ct = np.complex64 if prec == 32 else np.complex128
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    F = lambda v: ctypes.byref(R(v))
    reals = [np.ascontiguousarray(a, ft) for a in (vp, vs, rho, vpb, vsb, rhob)]
    flds = [np.ascontiguousarray(a, ct) for a in (ub, vb, wb)]
    sf = np.zeros(3 * nx * ny * nz, ct)
    ierr = ctypes.c_int(-1)
    L.wtref_hetfldsrc(I(nx), I(ny), I(nz), F(dx), F(freq), F(tau), I(int(fst)), I(int(usemin)), F(pmld),
                      F(pmlr), *[_ptr(a) for a in reals], *[_ptr(a) for a in flds], F(complex(srcfn).real),
                      F(complex(srcfn).imag), _ptr(sf), ctypes.byref(ierr))
    return ierr.value, sf


def ref_makeuh(nx, ny, nz, dx, freq, vp, vs, rho, sfld, ub, vb, wb, rec, hyd, fst=True, totfld=True, prec=64):
    """Reference MAKEUH (makeuh.f:1) through oracle/_ref -> (ierr, uest, vest, west, utest).

    rec: dict x, y, z, spread, ireg, fac (None: recin = 0); hyd: dict x, y, z, spread, ireg."""
    L = ref(prec)
    R = ctypes.c_float if prec == 32 else ctypes.c_double
    ft = np.float32 if prec == 32 else np.float64
    ct = np.complex64 if prec == 32 else np.complex128
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    F = lambda v: ctypes.byref(R(v))
    A = lambda v, t: np.ascontiguousarray(v, t)
    ng, nr = len(rec["x"]), len(hyd["x"])
    xg, yg, zg = (A(rec[k], ft) for k in ("x", "y", "z"))
    xr, yr, zr = (A(hyd[k], ft) for k in ("x", "y", "z"))
    fac = A(rec["fac"] if rec.get("fac") is not None else np.ones(max(ng, 1)), ct)
    reals = [A(a, ft) for a in (vp, vs, rho)]
    flds = [A(a, ct) for a in (sfld, ub, vb, wb)]
    ue, ve, we = (np.zeros(max(ng, 1), ct) for _ in range(3))
    pe = np.zeros(max(nr, 1), ct)
    ierr = ctypes.c_int(-1)
    L.wtref_makeuh(I(nx), I(ny), I(nz), F(dx), F(freq), I(int(fst)), *[_ptr(a) for a in reals], I(int(totfld)),
                   *[_ptr(a) for a in flds], I(ng), _ptr(xg), _ptr(yg), _ptr(zg), F(rec["spread"]),
                   I(rec["ireg"]), I(0 if rec.get("fac") is None else 1), _ptr(fac), I(nr), _ptr(xr), _ptr(yr),
                   _ptr(zr), F(hyd["spread"]), I(hyd["ireg"]), _ptr(ue), _ptr(ve), _ptr(we), _ptr(pe),
                   ctypes.byref(ierr))
    return ierr.value, ue[:ng], ve[:ng], we[:ng], pe[:nr]


def ref_dispersion(vp, vs, qpi, qsi, rho, freqbase, omega_re, prec=64):
    """Reference DISPERSION (dispersion.f:1) -> (lambda, mu); arrays (ny, nx, nz)."""
    L = ref(prec)
    R = ctypes.c_float if prec == 32 else ctypes.c_double
    ft = np.float32 if prec == 32 else np.float64
    ct = np.complex64 if prec == 32 else np.complex128
    ny, nx, nz = np.shape(vp)
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    F = lambda v: ctypes.byref(R(v))
    arrs = [np.ascontiguousarray(a, ft) for a in (vp, vs, qpi, qsi, rho)]
    lam = np.zeros((ny, nx, nz), ct)
    mu = np.zeros((ny, nx, nz), ct)
    L.wtref_dispersion(I(nx), I(ny), I(nz), *[_ptr(a) for a in arrs], F(freqbase), F(omega_re), _ptr(lam), _ptr(mu))
    return lam, mu
