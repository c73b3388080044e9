"""Multifrontal solver handle: the host side of the reference's solver shim.

The reference drives STRUMPACK through ``strumpack_serial_finter`` jobs
1 (reorder), 2 (factor) and 3 (solve) (strumpack_finter.c:9-120); this class
exposes the same phases on one grid, for whole blocks of right-hand sides:

    s = Solver(nx, ny, nz)                  # job 1: ordering + symbolic
    s.set_model(lam, mu, rho)               # MODEL_MODULE state
    s.assemble(params)                      # ASMBLE at one frequency (FP64)
    s.factor(32)                            # job 2, complex64 fronts
    x, st = s.solve(b, prec=32, tol=1e-8)   # job 3 for a block, FP64 refinement
"""
import ctypes

import numpy as np

from ._lib import lib, check, WtSolveStats, WtPivotReport, NotConverged, WT_ERR_NOCONV

_vp = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(_vp)


class Solver:
    def __init__(self, nx, ny, nz, leaf_nodes=64):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = 3 * nx * ny * nz
        h = ctypes.c_void_p()
        check(lib().wt_solver_create(nx, ny, nz, leaf_nodes, ctypes.byref(h)), "wt_solver_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().wt_solver_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        arr = (ctypes.c_int64 * 8)()
        fl = ctypes.c_double()
        check(lib().wt_solver_info(self._h, arr, ctypes.byref(fl)), "wt_solver_info")
        keys = ("n", "nnz", "fronts", "levels", "max_p", "max_u", "factor_entries", "cb_entries")
        d = dict(zip(keys, list(arr)))
        d["flops"] = fl.value
        return d

    def bytes_needed(self, prec=32, nrhs=64):
        return int(lib().wt_solver_bytes(self._h, prec, nrhs))

    def set_model(self, lam, mu, rho):
        lam = np.ascontiguousarray(lam, np.complex128).ravel()
        mu = np.ascontiguousarray(mu, np.complex128).ravel()
        rho = np.ascontiguousarray(rho, np.float64).ravel()
        assert lam.size == self.n // 3 and mu.size == lam.size and rho.size == lam.size
        check(lib().wt_solver_set_model(self._h, _p(lam), _p(mu), _p(rho), 0), "wt_solver_set_model")

    def set_model_q(self, vp, vs, qpi, qsi, rho, freqbase, omega):
        """MODEL_MODULE state of an attenuating model at `omega`: the complex
        moduli of DISPERSION (dispersion.f:1-131) formed on the device from
        speeds at `freqbase` Hz and 1/Qp, 1/Qs (arrays z fastest)."""
        arrs = [np.ascontiguousarray(np.asarray(a, np.float64).ravel()) for a in (vp, vs, qpi, qsi, rho)]
        check(lib().wt_solver_set_model_q(self._h, *[_p(a) for a in arrs], float(freqbase), float(np.real(omega))),
              "wt_solver_set_model_q")

    def assemble(self, params):
        check(lib().wt_solver_assemble(self._h, ctypes.byref(params)), "wt_solver_assemble")

    def set_values(self, vals):
        vals = np.ascontiguousarray(vals)
        prec = 32 if vals.dtype == np.complex64 else 64
        vals = vals.astype(np.complex64 if prec == 32 else np.complex128, copy=False)
        check(lib().wt_solver_set_values(self._h, _p(vals), prec, 0), "wt_solver_set_values")

    def values(self):
        from .assembly import graph_nnz
        v = np.zeros(graph_nnz(self.nx, self.ny, self.nz), np.complex128)
        check(lib().wt_solver_get_values(self._h, _p(v)), "wt_solver_get_values")
        return v

    def apply(self, x):
        """A x with the FP64 CSR values (x: (n,) or (n, nrhs))."""
        x = np.asarray(x)
        one = x.ndim == 1
        X = np.ascontiguousarray(x.reshape(self.n, -1), np.complex128)
        Y = np.zeros_like(X)
        check(lib().wt_solver_apply(self._h, X.shape[1], _p(X), _p(Y), 0), "wt_solver_apply")
        return Y[:, 0] if one else Y

    def factor(self, prec=32):
        check(lib().wt_solver_factor(self._h, prec), "wt_solver_factor")

    def pivot_report(self, prec=32):
        """Pivot statistics of the last factorization in ``prec``: zero_pivots,
        tiny_pivots, min_pivot_ratio and max_growth (``wt_pivot_report``)."""
        r = WtPivotReport()
        check(lib().wt_solver_pivot_report(self._h, prec, ctypes.byref(r)), "wt_solver_pivot_report")
        return {k: getattr(r, k) for k, _ in WtPivotReport._fields_}

    def autotune(self, prec=32):
        """Times every factorization GEMM launch under each tile variant and
        keeps the fastest; returns the GEMM speedup over the default tiles."""
        sp = ctypes.c_double()
        check(lib().wt_solver_autotune(self._h, prec, ctypes.byref(sp)), "wt_solver_autotune")
        return sp.value

    def set_gemm_split(self, enable=True):
        """complex64 GEMMs as three-way bf16 splits on the bf16 matrix cores
        (default), or all on the fp32 matrix cores."""
        check(lib().wt_solver_set_gemm_split(self._h, int(bool(enable))), "wt_solver_set_gemm_split")

    def set_splitk(self, min_depth=512):
        """Split the back substitution's long-K update products into pieces
        of at least ``min_depth`` (0: never); from the next solve."""
        check(lib().wt_solver_set_splitk(self._h, int(min_depth)), "wt_solver_set_splitk")

    def set_graphs(self, enable=True):
        """Capture the factorization in a HIP graph and replay it."""
        check(lib().wt_solver_set_graphs(self._h, int(bool(enable))), "wt_solver_set_graphs")

    def get_tiles(self, prec=32, which=0, nrhs=64):
        """Tile variant of every GEMM launch of the factorization (which=0)
        or of the block solve for nrhs columns (which=1)."""
        n = ctypes.c_int64()
        check(lib().wt_solver_get_tiles(self._h, prec, which, nrhs, None, ctypes.byref(n)), "wt_solver_get_tiles")
        v = np.zeros(n.value, np.int32)
        check(lib().wt_solver_get_tiles(self._h, prec, which, nrhs, _p(v), ctypes.byref(n)), "wt_solver_get_tiles")
        return v

    def set_tiles(self, vars, prec=32, which=0, nrhs=64):
        v = np.ascontiguousarray(vars, np.int32)
        check(lib().wt_solver_set_tiles(self._h, prec, which, nrhs, _p(v), v.size), "wt_solver_set_tiles")

    def release(self, prec):
        check(lib().wt_solver_release(self._h, prec), "wt_solver_release")

    PROFILE_CLASSES = ("gemm_factor", "gemm_solve", "diag_tile", "extend_add", "fill_scatter",
                       "assemble", "residual", "reorder", "panel_trsm", "gemm_factor_narrow", "gemm_factor_split")

    def profile(self, enable=True):
        """Enable (and reset) per-kernel HIP-event timing on the solver stream."""
        check(lib().wt_solver_profile(self._h, int(enable)), "wt_solver_profile")

    def profile_read(self):
        out = {}
        for i, name in enumerate(self.PROFILE_CLASSES):
            n = ctypes.c_int64()
            ms, fl, by = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            check(lib().wt_solver_profile_read(self._h, i, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(fl),
                                               ctypes.byref(by)), "wt_solver_profile_read")
            out[name] = dict(launches=n.value, ms=ms.value, flops=fl.value, bytes=by.value)
        return out

    def solve(self, b, prec=32, tol=1e-8, max_iter=20):
        """b: (n,) or (n, nrhs) complex; returns (x, stats dict).  Raises
        NotConverged (with .x and .stats) when refinement ends above tol;
        tol <= 0 runs exactly max_iter solves."""
        b = np.asarray(b)
        one = b.ndim == 1
        B = np.ascontiguousarray(b.reshape(self.n, -1), np.complex128)
        X = np.zeros_like(B)
        st = WtSolveStats()
        rc = lib().wt_solver_solve(self._h, prec, B.shape[1], _p(B), _p(X), 0, tol, max_iter, ctypes.byref(st))
        stats = {k: getattr(st, k) for k, _ in WtSolveStats._fields_}
        x = X[:, 0] if one else X
        if rc == WT_ERR_NOCONV:
            raise NotConverged("wt_solver_solve", stats, x)
        check(rc, "wt_solver_solve")
        return x, stats

    def stream(self):
        """The HIP stream of this handle (a ctypes void pointer)."""
        st = ctypes.c_void_p()
        check(lib().wt_solver_stream(self._h, ctypes.byref(st)), "wt_solver_stream")
        return st

    def scattered_source(self, lam_bg, mu_bg, rho_bg, ubg, srcfn=None, out=None, nrhs=None):
        """HETFLDSRC (hetfldsrc.f90:17-1039): srcfn_j (A_bg - A) u_bg_j with A
        the operator of the last ``assemble`` and A_bg the background model's at
        the same parameters.  ubg: a device block (rhs.DeviceBuffer, row-major
        (n, nrhs) complex128) or a host array (n, nrhs); returns the same kind
        (into ``out``, a DeviceBuffer, when given).  For a device ubg the
        column count is ``nrhs`` (default: all the buffer holds)."""
        from .rhs import DeviceBuffer
        host = not isinstance(ubg, DeviceBuffer)
        if host:
            ubg = np.ascontiguousarray(np.asarray(ubg).reshape(self.n, -1), np.complex128)
            nrhs = ubg.shape[1]
            ub = DeviceBuffer.from_host(ubg)
        else:
            nrhs = int(nrhs) if nrhs is not None else ubg.nbytes // (self.n * 16)
            ub = ubg
        lb = np.ascontiguousarray(lam_bg, np.complex128).ravel()
        mb = np.ascontiguousarray(mu_bg, np.complex128).ravel()
        rb = np.ascontiguousarray(rho_bg, np.float64).ravel()
        sf = None
        if srcfn is not None:
            sf = np.ascontiguousarray(np.broadcast_to(np.asarray(srcfn, np.complex128), (nrhs,)))
        if out is None or out.nbytes < self.n * nrhs * 16:
            out = DeviceBuffer(self.n * nrhs * 16)
        check(lib().wt_solver_scattered_source(self._h, _p(lb), _p(mb), _p(rb), 0, nrhs, ub.ptr,
                                               None if sf is None else _p(sf), out.ptr),
              "wt_solver_scattered_source")
        if host:
            return out.to_host(np.complex128, (self.n, nrhs))
        return out
