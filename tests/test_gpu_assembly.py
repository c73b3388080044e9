"""GPU assembly parity: the HIP kernels against the oracle (bit-exact) and
against the reference's golden vectors (tolerances as stated)."""
import ctypes
import hashlib

import numpy as np
import pytest

import oracle_lib as O
import wtomo3d_amd as W
from helpers import cases, load, ref_params, lame_ref, rel_err_entries

pytestmark = pytest.mark.gpu
CASE_NAMES = [k for k, v in cases().items() if isinstance(v, dict) and "seed" in v]


def _params(nx, ny, nz, dx, om, pmld, pmlf, ipml, fst=True, usemin=True):
    return W._lib.WtGridParams(nx, ny, nz, ipml, int(fst), int(usemin), dx, pmld, pmlf, om.real, om.imag)


@pytest.mark.parametrize("name", CASE_NAMES)
def test_gpu_pattern_bit_exact(name):
    m = cases()[name]
    g = load(name)
    ir, jc = W.graph(m["nx"], m["ny"], m["nz"], base=1)
    assert np.array_equal(ir, g["irptr"]) and np.array_equal(jc, g["jcptr"])


def test_gpu_pattern_c1_digest_and_base0():
    d = cases()["c1_pattern_sha256"]
    ir, jc = W.graph(24, 24, 24, base=1)
    assert hashlib.sha256(ir.tobytes()).hexdigest() == d["irptr"]
    assert hashlib.sha256(jc.tobytes()).hexdigest() == d["jcptr"]
    ir0, jc0 = W.graph(24, 24, 24, base=0)
    assert np.array_equal(ir0, ir - 1) and np.array_equal(jc0, jc - 1)


@pytest.mark.parametrize("dims", [(96, 96, 64), (17, 40, 23)])
def test_gpu_pattern_large_equals_oracle(dims):
    ir, jc = W.graph(*dims, base=1)
    ir2, jc2 = O.graph(*dims)
    assert np.array_equal(ir, ir2) and np.array_equal(jc, jc2)


def _gpu_case(name, prec):
    m = cases()[name]
    g = load(name)
    dx, om, pmld, pmlf, ipml = ref_params(m, prec)
    ft = np.float32 if prec == 32 else np.float64
    lam, mu, rho = lame_ref(g["vp"].astype(ft), g["vs"].astype(ft), g["rho"].astype(ft), prec)
    usemin = bool(m.get("usemin", 1))
    p = _params(m["nx"], m["ny"], m["nz"], dx, om, pmld, pmlf, ipml, bool(m["fst"]), usemin)
    a = W.assemble(p, lam, mu, rho, prec=prec)
    rc, o = O.assemble(m["nx"], m["ny"], m["nz"], dx, om, lam, mu, rho, pmld, pmlf, ipml,
                       fst=bool(m["fst"]), usemin=usemin, prec=prec)
    assert rc == 0
    return a, o, g["vals%d" % prec]


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("name", CASE_NAMES)
def test_gpu_values_vs_oracle_and_reference(name, prec):
    a, o, ref = _gpu_case(name, prec)
    # the kernel and the oracle evaluate the same operations in the same
    # order, but the kernel's complex products may fuse into multiply-adds
    # (the oracle is compiled without contraction): a few ulps, amplified
    # where stretched PML terms cancel (observed <= 6.1e-15 in FP64)
    assert np.array_equal(a.view(np.uint8), o.view(np.uint8)) or \
        rel_err_entries(a, o)[0] <= (1e-14 if prec == 64 else 4e-6)
    rel, zm = rel_err_entries(a, ref)
    assert zm == 0
    assert rel <= (1e-12 if prec == 64 else 3e-5)


@pytest.mark.parametrize("prec", [32, 64])
def test_gpu_values_layered_64x64x48(prec):
    """The layered 64x64x48 configuration (dx = 250 m, 0.25 Hz), whole matrix."""
    nx, ny, nz, dx = 64, 64, 48, 250.0
    rng = np.random.default_rng(1002)
    z = np.arange(nz) * dx
    tops = np.sort(rng.uniform(0, nz * dx, 5))
    vps = np.sort(rng.uniform(4000, 8000, 6))
    vp1 = vps[np.searchsorted(tops, z, side="right")]
    ft = np.float32 if prec == 32 else np.float64
    vp = np.broadcast_to(vp1, (ny, nx, nz)).astype(ft)
    vs = (vp / ft(1.73)).astype(ft)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(ft)
    meta = dict(dx=dx, freq=0.25, pmld=10 * dx, pmlr=1e-3)
    dx_, om, pmld, pmlf, ipml = ref_params(meta, prec)
    lam, mu, rho1 = lame_ref(vp, vs, rho, prec)
    a = W.assemble(_params(nx, ny, nz, dx_, om, pmld, pmlf, ipml), lam, mu, rho1, prec=prec)
    rc, o = O.assemble(nx, ny, nz, dx_, om, lam, mu, rho1, pmld, pmlf, ipml, prec=prec)
    assert rc == 0
    assert np.array_equal(a.view(np.uint8), o.view(np.uint8)) or \
        rel_err_entries(a, o)[0] <= (1e-15 if prec == 64 else 4e-6)


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("dims", [(6, 5, 150), (4, 3, 65), (5, 4, 128), (5, 3, 16), (7, 6, 17), (4, 6, 49)])
@pytest.mark.parametrize("usemin", [True, False])
def test_gpu_values_several_z_segments(dims, prec, usemin):
    """Columns longer than one 64-node segment (a one-node last segment, an
    exact multiple, three segments) and columns ending at, one past and one
    past three 16-node output chunks: every chunk's rows land at their CSR
    offsets; heterogeneous random model, PML on every side, with the
    Min-cell (usemin) and the traction-free top."""
    nx, ny, nz = dims
    rng = np.random.default_rng(77 + nz)
    ft = np.float32 if prec == 32 else np.float64
    vp = (4500 + 2000 * rng.random((ny, nx, nz))).astype(ft)
    vs = (vp / ft(1.8)).astype(ft)
    rho = (2000 + 800 * rng.random((ny, nx, nz))).astype(ft)
    meta = dict(dx=250.0, freq=0.3, pmld=500.0, pmlr=1e-3)
    dx_, om, pmld, pmlf, ipml = ref_params(meta, prec)
    lam, mu, rho1 = lame_ref(vp, vs, rho, prec)
    a = W.assemble(_params(nx, ny, nz, dx_, om, pmld, pmlf, ipml, True, usemin), lam, mu, rho1, prec=prec)
    rc, o = O.assemble(nx, ny, nz, dx_, om, lam, mu, rho1, pmld, pmlf, ipml, usemin=usemin, prec=prec)
    assert rc == 0
    # a misplaced entry would be off by O(1); rounding of the random model's
    # cell averages stays at the assembly bar (1e-12 FP64, 3e-5 FP32)
    assert rel_err_entries(a, o)[0] <= (1e-12 if prec == 64 else 3e-5)


def test_gpu_c1_sampled_rows_vs_reference():
    s = load("c1_rows")
    nx = ny = nz = 24
    ir, _ = O.graph(nx, ny, nz)
    idx = np.concatenate([np.arange(ir[r] - 1, ir[r + 1] - 1) for r in s["rows"]])
    meta = dict(dx=500.0, freq=0.2, pmld=5000.0, pmlr=1e-3)
    for prec, tol in ((64, 1e-12), (32, 3e-5)):
        ft = np.float32 if prec == 32 else np.float64
        vp = np.full((ny, nx, nz), 6000.0, ft)
        vs = (vp / ft(np.sqrt(3.0))).astype(ft)
        rho = np.full_like(vp, 2700.0)
        dx, om, pmld, pmlf, ipml = ref_params(meta, prec)
        lam, mu, rho1 = lame_ref(vp, vs, rho, prec)
        a = W.assemble(_params(nx, ny, nz, dx, om, pmld, pmlf, ipml), lam, mu, rho1, prec=prec)
        rel, zm = rel_err_entries(a[idx], s["vals%d" % prec])
        assert zm == 0 and rel <= tol


def test_gpu_asmble_dropin_entry():
    """wt_asmble: the reference's ASMBLE argument list + module state."""
    name = "g667_fs"
    m = cases()[name]
    g = load(name)
    dx, om, pmld, pmlf, ipml = ref_params(m, 32)
    lam, mu, rho = lame_ref(g["vp"].astype(np.float32), g["vs"].astype(np.float32),
                            g["rho"].astype(np.float32), 32)
    ir, jc = g["irptr"], g["jcptr"]
    a = np.zeros(jc.size, np.complex64)
    ierr = ctypes.c_int(-1)
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    L = W.lib()
    L.wt_asmble(6, jc.size, m["nx"], m["ny"], m["nz"], P(ir), P(jc), P(a), ctypes.byref(ierr),
                P(lam), P(mu), P(rho), om.real, om.imag, dx, pmld, pmlf, ipml, 1, 1)
    assert ierr.value == 0
    rel, zm = rel_err_entries(a, g["vals32"])
    assert zm == 0 and rel <= 3e-5
    # a corrupted row pointer is reported like asmble.f90:556-564
    bad = ir.copy()
    bad[5] += 1
    L.wt_asmble(6, jc.size, m["nx"], m["ny"], m["nz"], P(bad), P(jc), P(a), ctypes.byref(ierr),
                P(lam), P(mu), P(rho), om.real, om.imag, dx, pmld, pmlf, ipml, 1, 1)
    assert ierr.value != 0
    # wrong predicted nonzero count (asmble.f90:619-626)
    L.wt_asmble(6, jc.size - 3, m["nx"], m["ny"], m["nz"], P(ir), P(jc), P(a), ctypes.byref(ierr),
                P(lam), P(mu), P(rho), om.real, om.imag, dx, pmld, pmlf, ipml, 1, 1)
    assert ierr.value != 0


@pytest.mark.skipif(not O.have_ref(), reason="reference build absent")
@pytest.mark.parametrize("prec", [64, 32])
def test_gpu_values_heterogeneous_96x96x64_vs_reference(prec):
    """The benchmark's operator, whole matrix: the heterogeneous 96x96x64 model
    (dx = 200 m, 0.3 Hz, 10-node PML, free surface) assembled on the GPU
    against the reference's own ASMBLE compiled from its sources (FP64 build
    for the FP64 bar)."""
    from wtomo3d_amd import media
    nx, ny, nz = 96, 96, 64
    meta = dict(dx=200.0, freq=0.3, pmld=2000.0, pmlr=1e-3)
    ft = np.float32 if prec == 32 else np.float64
    vp, vs, rho = (a.astype(ft) for a in media.heterogeneous(nx, ny, nz, meta["dx"], seed=1003))
    dx_, om, pmld, pmlf, ipml = ref_params(meta, prec)
    lam, mu, rho1 = lame_ref(vp, vs, rho, prec)
    a = W.assemble(_params(nx, ny, nz, dx_, om, pmld, pmlf, ipml), lam, mu, rho1, prec=prec)
    ierr, ir, jc, ref = O.ref_assemble(nx, ny, nz, meta["dx"], meta["freq"], vp, vs, rho, meta["pmld"],
                                       meta["pmlr"], prec=prec)
    assert ierr == 0 and ref.size == a.size
    rel, zm = rel_err_entries(a, ref)
    assert zm == 0
    assert rel <= (1e-12 if prec == 64 else 3e-5), rel


@pytest.mark.skipif(not O.have_ref(), reason="reference build absent")
@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("tau", [4.0, 0.8])
def test_gpu_values_complex_frequency_vs_reference(prec, tau):
    """The anti-aliasing damping (tau set): the GPU assembly at the complex
    frequency omega = 2 pi f + i / tau against the reference's ASMBLE, which
    forms it at elwave.f90:164, on a layered + random model with PML on every
    side and the free surface."""
    nx, ny, nz = 22, 19, 17
    meta = dict(dx=250.0, freq=0.25, pmld=1250.0, pmlr=1e-3, tau=tau)
    ft = np.float32 if prec == 32 else np.float64
    rng = np.random.default_rng(91)
    vp = (4000 + 3000 * rng.random((ny, nx, nz))).astype(ft)
    vs = (vp / ft(1.8)).astype(ft)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(ft)
    dx_, om, pmld, pmlf, ipml = ref_params(meta, prec)
    assert om.imag != 0.0
    lam, mu, rho1 = lame_ref(vp, vs, rho, prec)
    a = W.assemble(_params(nx, ny, nz, dx_, om, pmld, pmlf, ipml), lam, mu, rho1, prec=prec)
    ierr, ir, jc, ref = O.ref_assemble(nx, ny, nz, meta["dx"], meta["freq"], vp, vs, rho, meta["pmld"],
                                       meta["pmlr"], tau=tau, prec=prec)
    assert ierr == 0 and ref.size == a.size
    rel, zm = rel_err_entries(a, ref)
    assert zm == 0
    assert rel <= (1e-12 if prec == 64 else 3e-5), rel
