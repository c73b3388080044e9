"""Shared test helpers: golden fixtures and reference-faithful parameters."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
PI = 3.141592653589793


def cases():
    with open(os.path.join(GOLDEN, "assembly_cases.json")) as f:
        return json.load(f)


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def ref_params(meta, prec):
    """(dx, omega, pmld, pmlf, ipml) exactly as the reference forms them in
    precision ``prec``: readini.f90:81-84 and elwave.f90:151,164."""
    ft = np.float32 if prec == 32 else np.float64
    dx, pmld, pmlr = ft(meta["dx"]), ft(meta["pmld"]), ft(meta["pmlr"])
    ipml = int(pmld / dx) + 1
    pmlf = ft(3.0) * ft(np.log(ft(1.0) / pmlr)) / (ft(2.0) * pmld * pmld * pmld)
    om_re = ft(2.0) * ft(PI) * ft(meta["freq"])
    om_im = ft(0.0)
    if meta.get("tau", 999.999) != 999.999:
        om_im = ft(1.0) / ft(meta["tau"])
    return float(dx), complex(float(om_re), float(om_im)), float(pmld), float(pmlf), ipml


def lame_ref(vp, vs, rho, prec):
    """modelio.f90:78-92 in precision ``prec`` (imaginary parts are zero)."""
    ft = np.float32 if prec == 32 else np.float64
    ct = np.complex64 if prec == 32 else np.complex128
    vp, vs, rho = (np.asarray(a, ft) for a in (vp, vs, rho))
    mu = (vs * vs) * rho
    lam = (vp * vp) * rho - ft(2.0) * mu
    return lam.astype(ct).ravel(), mu.astype(ct).ravel(), rho.ravel()


def rel_err_entries(a, ref):
    """max_i |a_i - ref_i| / |ref_i| over nonzero ref, and the count of
    entries where ref is exactly zero but a is not."""
    a = np.asarray(a)
    ref = np.asarray(ref)
    s = np.abs(ref)
    nz = s > 0
    rel = np.abs(a[nz] - ref[nz]) / s[nz]
    return (float(rel.max()) if rel.size else 0.0), int(np.count_nonzero(a[~nz]))
