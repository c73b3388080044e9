"""Loader for libwtomo3d_hip.so and the C structs it shares with Python."""
import ctypes
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libwtomo3d_hip.so")


class LibraryMissing(RuntimeError):
    pass


# return codes of include/wtomo3d.h
WT_OK, WT_ERR_ARG, WT_ERR_HIP, WT_ERR_STATE, WT_ERR_UNSUPPORTED = 0, 1, 2, 3, 4
WT_ERR_NOMEM, WT_ERR_PATTERN, WT_ERR_PIVOT, WT_ERR_NOCONV = 5, 6, 7, 8


class WtError(RuntimeError):
    """A library call returned a WT_ERR_* code (``.code``)."""

    def __init__(self, what, code):
        super().__init__("%s failed with code %d" % (what, code))
        self.what, self.code = what, code


class NotConverged(WtError):
    """Refinement ended above its tolerance (WT_ERR_NOCONV).  ``stats`` (a
    dict, or a list of dicts for a batch) and ``x`` (the solution reached,
    when the caller's arrays were on the host) are attached."""

    def __init__(self, what, stats=None, x=None):
        super().__init__(what, WT_ERR_NOCONV)
        self.stats, self.x = stats, x


class WtGridParams(ctypes.Structure):
    """Mirror of ``wt_grid_params`` (include/wtomo3d.h)."""
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("ipml", ctypes.c_int32), ("fst", ctypes.c_int32), ("usemin", ctypes.c_int32),
                ("dx", ctypes.c_double), ("pmld", ctypes.c_double), ("pmlf", ctypes.c_double),
                ("omega_re", ctypes.c_double), ("omega_im", ctypes.c_double)]


class WtSolveStats(ctypes.Structure):
    """Mirror of ``wt_solve_stats``."""
    _fields_ = [("iterations", ctypes.c_int32), ("nrhs", ctypes.c_int32),
                ("rel_residual", ctypes.c_double), ("rel_residual_first", ctypes.c_double),
                ("abs_residual", ctypes.c_double), ("ms", ctypes.c_double),
                ("converged", ctypes.c_int32), ("stalled", ctypes.c_int32)]


class WtPivotReport(ctypes.Structure):
    """Mirror of ``wt_pivot_report``."""
    _fields_ = [("zero_pivots", ctypes.c_int32), ("tiny_pivots", ctypes.c_int32),
                ("min_pivot_ratio", ctypes.c_double), ("max_growth", ctypes.c_double)]


_LIB = None
_vp = ctypes.c_void_p


def _declare(L):
    i32, i64 = ctypes.c_int, ctypes.c_int64
    L.wt_graph_nnz.argtypes = [i32] * 3
    L.wt_graph_nnz.restype = i64
    L.wt_graph.argtypes = [i32] * 4 + [_vp, _vp]
    L.wt_graph_dev.argtypes = [i32] * 4 + [_vp, _vp, _vp]
    L.wt_assemble.argtypes = [ctypes.POINTER(WtGridParams), i32, _vp, _vp, _vp, _vp]
    L.wt_assemble_dev.argtypes = [ctypes.POINTER(WtGridParams), i32, _vp, _vp, _vp, _vp, _vp, _vp]
    L.wt_asmble.argtypes = [i32] * 5 + [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_int), _vp, _vp, _vp] + \
        [ctypes.c_float] * 5 + [i32] * 3
    L.wt_asmble.restype = None
    L.wt_dev_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]
    L.wt_dev_free.argtypes = [_vp]
    L.wt_h2d.argtypes = [_vp, _vp, ctypes.c_size_t]
    L.wt_d2h.argtypes = [_vp, _vp, ctypes.c_size_t]
    L.wt_dev_sync.argtypes = []
    L.wt_device_count.argtypes = []
    for name in ("wt_graph", "wt_graph_dev", "wt_assemble", "wt_assemble_dev", "wt_dev_alloc",
                 "wt_dev_free", "wt_h2d", "wt_d2h", "wt_dev_sync", "wt_device_count"):
        getattr(L, name).restype = ctypes.c_int
    i32p = ctypes.POINTER(ctypes.c_int)
    L.wt_solver_create.argtypes = [i32] * 4 + [ctypes.POINTER(ctypes.c_void_p)]
    L.wt_solver_destroy.argtypes = [_vp]
    L.wt_solver_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)]
    L.wt_solver_bytes.argtypes = [_vp, i32, i32]
    L.wt_solver_bytes.restype = i64
    L.wt_solver_set_model.argtypes = [_vp, _vp, _vp, _vp, i32]
    L.wt_solver_assemble.argtypes = [_vp, ctypes.POINTER(WtGridParams)]
    L.wt_solver_set_values.argtypes = [_vp, _vp, i32, i32]
    L.wt_solver_get_values.argtypes = [_vp, _vp]
    L.wt_solver_factor.argtypes = [_vp, i32]
    L.wt_solver_release.argtypes = [_vp, i32]
    L.wt_solver_solve.argtypes = [_vp, i32, i32, _vp, _vp, i32, ctypes.c_double, i32,
                                  ctypes.POINTER(WtSolveStats)]
    for name in ("wt_solver_create", "wt_solver_destroy", "wt_solver_info", "wt_solver_set_model",
                 "wt_solver_assemble", "wt_solver_set_values", "wt_solver_get_values", "wt_solver_factor",
                 "wt_solver_release", "wt_solver_solve"):
        getattr(L, name).restype = ctypes.c_int
    L.wt_gemm_host.argtypes = [i32] * 4 + [_vp, i64, _vp, i64, _vp, i64, i32, _vp, i32,
                                           ctypes.POINTER(ctypes.c_double)]
    L.wt_gemm_host.restype = ctypes.c_int
    L.wt_gemm_host_var.argtypes = [i32] * 5 + [_vp, i64, _vp, i64, _vp, i64, i32, _vp, i32,
                                               ctypes.POINTER(ctypes.c_double)]
    L.wt_gemm_host_var.restype = ctypes.c_int
    pstats = ctypes.POINTER(WtSolveStats)
    L.wt_batch_create.argtypes = [i32, _vp, i32, i64, i64, ctypes.POINTER(ctypes.c_void_p)]
    L.wt_batch_destroy.argtypes = [_vp]
    L.wt_batch_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)]
    L.wt_batch_offsets.argtypes = [_vp, _vp, _vp]
    L.wt_batch_set_graphs.argtypes = [_vp, i32]
    L.wt_batch_set_models.argtypes = [_vp, _vp, _vp, _vp, i32]
    L.wt_batch_assemble.argtypes = [_vp, ctypes.POINTER(WtGridParams)]
    L.wt_batch_factor.argtypes = [_vp, i32]
    L.wt_batch_autotune.argtypes = [_vp, i32, ctypes.POINTER(ctypes.c_double)]
    L.wt_batch_solve.argtypes = [_vp, i32, i32, _vp, _vp, _vp, i32, ctypes.c_double, i32, pstats]
    L.wt_batch_apply.argtypes = [_vp, i32, _vp, _vp, i32]
    L.wt_batch_run.argtypes = [_vp, i32, i32, _vp, _vp, _vp, i32, ctypes.c_double, i32, pstats]
    L.wt_solver_autotune.argtypes = [_vp, i32, ctypes.POINTER(ctypes.c_double)]
    L.wt_solver_autotune.restype = ctypes.c_int
    L.wt_solver_pivot_report.argtypes = [_vp, i32, ctypes.POINTER(WtPivotReport)]
    L.wt_solver_pivot_report.restype = ctypes.c_int
    L.wt_solver_set_gemm_split.argtypes = [_vp, i32]
    L.wt_solver_set_gemm_split.restype = ctypes.c_int
    L.wt_solver_set_splitk.argtypes = [_vp, i32]
    L.wt_solver_set_splitk.restype = ctypes.c_int
    L.wt_solver_get_tiles.argtypes = [_vp, i32, i32, i32, _vp, ctypes.POINTER(ctypes.c_int64)]
    L.wt_solver_get_tiles.restype = ctypes.c_int
    L.wt_solver_set_tiles.argtypes = [_vp, i32, i32, i32, _vp, i64]
    L.wt_solver_set_tiles.restype = ctypes.c_int
    L.wt_solver_set_graphs.argtypes = [_vp, i32]
    L.wt_solver_set_graphs.restype = ctypes.c_int
    L.wt_batch_symbolic_info.argtypes = [i32, _vp, i32, ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_double)]
    L.wt_graph_host.argtypes = [i32] * 4 + [_vp, _vp]
    for name in ("wt_batch_create", "wt_batch_destroy", "wt_batch_info", "wt_batch_offsets", "wt_batch_set_graphs",
                 "wt_batch_set_models", "wt_batch_assemble", "wt_batch_factor", "wt_batch_autotune",
                 "wt_batch_solve", "wt_batch_apply", "wt_batch_symbolic_info", "wt_graph_host", "wt_batch_run"):
        getattr(L, name).restype = ctypes.c_int
    L.wt_solver_planewaves.argtypes = [_vp, i32, _vp, ctypes.c_double, ctypes.c_double, _vp]
    L.wt_solver_planewaves.restype = ctypes.c_int
    L.wt_estimate_bytes.argtypes = [i32] * 6
    L.wt_estimate_bytes.restype = i64
    L.wt_symbolic_info.argtypes = [i32] * 4 + [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)]
    L.wt_symbolic_info.restype = ctypes.c_int
    L.wt_dev_copy2d.argtypes = [_vp, ctypes.c_size_t, _vp] + [ctypes.c_size_t] * 3 + [_vp]
    L.wt_dev_copy2d.restype = ctypes.c_int
    L.wt_batch_stream.argtypes = [_vp, ctypes.POINTER(ctypes.c_void_p)]
    L.wt_batch_stream.restype = ctypes.c_int
    L.wt_dev_copy.argtypes = [_vp, _vp, ctypes.c_size_t]
    L.wt_dev_copy.restype = ctypes.c_int
    L.wt_solver_apply.argtypes = [_vp, i32, _vp, _vp, i32]
    L.wt_solver_apply.restype = ctypes.c_int
    L.wt_solver_profile.argtypes = [_vp, i32]
    L.wt_solver_profile.restype = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    L.wt_solver_profile_read.argtypes = [_vp, i32, ctypes.POINTER(ctypes.c_int64), dp, dp, dp]
    L.wt_solver_profile_read.restype = ctypes.c_int
    L.wt_haskgrn.argtypes = [i32, i32] + [_vp] * 6 + [i32, _vp, _vp, ctypes.c_double, i32, ctypes.c_double,
                                                     ctypes.c_double, _vp, _vp, _vp, _vp, _vp]
    L.wt_haskgrn.restype = ctypes.c_int
    L.wt_haskgrn_noalloc.argtypes = L.wt_haskgrn.argtypes
    L.wt_haskgrn_noalloc.restype = ctypes.c_int
    L.wt_planewave_layered.argtypes = [i32] * 3 + [ctypes.c_double, i32, i32, _vp, _vp, _vp, _vp, _vp, _vp]
    L.wt_planewave_layered.restype = ctypes.c_int
    L.wt_solver_scattered_source.argtypes = [_vp, _vp, _vp, _vp, i32, i32, _vp, _vp, _vp]
    L.wt_solver_scattered_source.restype = ctypes.c_int
    L.wt_scattered_source_grid.argtypes = [ctypes.POINTER(WtGridParams)] + [_vp] * 6 + [i32, i32, _vp, i64, _vp,
                                                                                _vp, i64, _vp]
    L.wt_scattered_source_grid.restype = ctypes.c_int
    L.wt_total_field.argtypes = [_vp, _vp, i64, _vp]
    L.wt_total_field.restype = ctypes.c_int
    L.wt_solver_stream.argtypes = [_vp, ctypes.POINTER(ctypes.c_void_p)]
    L.wt_solver_stream.restype = ctypes.c_int
    f64 = ctypes.c_double
    L.wt_receivers_create.argtypes = [i32] * 3 + [f64, _vp, i32, _vp, _vp, _vp, f64, i32, _vp, i32, _vp, _vp, _vp,
                                                  f64, i32, _vp, _vp, f64, f64, ctypes.POINTER(ctypes.c_void_p)]
    L.wt_receivers_create.restype = ctypes.c_int
    L.wt_receivers_rows.argtypes = [_vp]
    L.wt_receivers_rows.restype = ctypes.c_int
    L.wt_receivers_extract.argtypes = [_vp, i32, _vp, _vp, _vp]
    L.wt_receivers_extract.restype = ctypes.c_int
    L.wt_receivers_destroy.argtypes = [_vp]
    L.wt_receivers_destroy.restype = ctypes.c_int
    L.wt_dispersion.argtypes = [i64] + [_vp] * 5 + [f64, f64, _vp, _vp, i32, _vp]
    L.wt_dispersion.restype = ctypes.c_int
    L.wt_solver_set_model_q.argtypes = [_vp] + [_vp] * 5 + [f64, f64]
    L.wt_solver_set_model_q.restype = ctypes.c_int
    L.strumpack_serial_finter.argtypes = [i32p, i32p, i32p, _vp, _vp, _vp, _vp, _vp, i32p]
    L.strumpack_serial_finter.restype = None
    return L


def lib():
    """The loaded library; raises LibraryMissing when it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise LibraryMissing("%s is not built: run python -m wtomo3d_amd.build" % LIBPATH)
        _LIB = _declare(ctypes.CDLL(LIBPATH))
    return _LIB


def check(rc, what):
    if rc == WT_ERR_NOCONV:
        raise NotConverged(what)
    if rc != 0:
        raise WtError(what, rc)


def pml_params(dx, pmld=None, pmlr=1e-3, single=False):
    """(ipml, pmlf) as READINI_FORWARD computes them (readini.f90:81-84).

    ``single`` evaluates in float32 like the reference's default REAL."""
    import numpy as np
    if pmld is None:
        pmld = 10.0 * dx
    if single:
        f = np.float32
        dxf, pd, pr = f(dx), f(pmld), f(pmlr)
        ipml = int(pd / dxf) + 1
        pmlf = f(3.0) * f(np.log(f(1.0) / pr)) / (f(2.0) * pd * pd * pd)
        return ipml, float(pmlf)
    ipml = int(pmld / dx) + 1
    pmlf = 3.0 * math.log(1.0 / pmlr) / (2.0 * pmld ** 3)
    return ipml, pmlf


def grid_params(nx, ny, nz, dx, omega, pmld=None, pmlr=1e-3, fst=True, usemin=True, single=False):
    """Build a ``wt_grid_params`` for one frequency (omega may be complex)."""
    if pmld is None:
        pmld = 10.0 * dx
    ipml, pmlf = pml_params(dx, pmld, pmlr, single)
    om = complex(omega)
    return WtGridParams(nx, ny, nz, ipml, int(fst), int(usemin), float(dx), float(pmld), float(pmlf),
                        om.real, om.imag)
