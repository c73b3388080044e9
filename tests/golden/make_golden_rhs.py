"""Write the plane-wave right-hand-side golden fixtures from the compiled reference.

Runs only where oracle/_ref is built (the reference's own haskgrn.f and
hetfldsrc.f90 with the assembly sources, compiled by oracle/Makefile).

- haskgrn_cases.npz: per case the layer model, source and depth grid, and
  the reference's profiles (ugrn1f, wgrn1f) and ierr as compiled (complex64)
  and with 8-byte REAL (complex128).
- hetfldsrc_g877.npz: an 8x7x7 grid (ny = nz, where the reference's node
  numbering of hetfldsrc.f90:1020 is the unknown order) with a perturbed
  model, its layered background, background fields and srcfn, and the
  reference's scattered-field source in both precisions.
- makeuh_g121110.npz: a 12x11x10 grid with a scattered field and a
  background field, receivers (nearest node, Kaiser sinc, Dirac spread; at
  and just below the free surface so that taps are mirrored) and
  hydrophones, and the reference's MAKEUH outputs in both precisions.

    python tests/golden/make_golden_rhs.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle_lib as O  # noqa: E402
from wtomo3d_amd import media  # noqa: E402


def hk_cases():
    c3 = media.layer_stack(96, 96, 64, 200.0, 1003)
    thin = dict(top=[0.0, 900.0, 1000.0, 2600.0, 2700.0, 5000.0], vp=[3500.0, 4200.0, 5100.0, 5600.0, 6100.0, 7200.0],
                vs=[2000.0, 2400.0, 2950.0, 3250.0, 3500.0, 4150.0], rho=[2300.0, 2450.0, 2600.0, 2700.0, 2800.0, 3100.0])
    # a fluid lens between solid layers (a fluid at the free surface makes
    # the reference's surface determinant vanish: its profiles are NaN)
    fluid = dict(top=[0.0, 1500.0, 2400.0, 9000.0], vp=[3800.0, 1500.0, 6000.0, 7800.0],
                 vs=[2100.0, 0.0, 3450.0, 4400.0], rho=[2400.0, 1030.0, 2750.0, 3200.0])
    lvz = dict(top=[0.0, 2000.0, 5000.0, 7000.0, 12000.0], vp=[5000.0, 6500.0, 4800.0, 6800.0, 8000.0],
               vs=[2900.0, 3750.0, 2700.0, 3900.0, 4600.0], rho=[2500.0, 2800.0, 2450.0, 2900.0, 3300.0])
    for d in (thin, fluid, lvz):
        d["vpd"] = [v * 0.9995 for v in d["vp"]]
        d["vsd"] = [v * 0.999 for v in d["vs"]]
    w3 = 2 * np.pi * 0.3
    # (name, layers, ain, is_s, omega, nz, dz, zorig)
    return [
        ("c3_p20", c3, np.deg2rad(20.0), 0, w3, 64, 200.0, 0.0),
        ("c3_s50_evanescent", c3, np.deg2rad(50.0), 1, w3, 64, 200.0, 0.0),
        ("thin_layers_skipped_s", thin, np.deg2rad(25.0), 1, 2 * np.pi * 0.45, 40, 250.0, 0.0),
        ("fluid_lens_p", fluid, np.deg2rad(15.0), 0, 2 * np.pi * 0.2, 50, 300.0, 0.0),
        ("lvz_s", lvz, np.deg2rad(35.0), 1, 2 * np.pi * 0.35, 70, 220.0, 0.0),
        ("above_top_error", lvz, np.deg2rad(20.0), 0, w3, 20, 200.0, -700.0),
    ]


def main():
    if not O.have_ref():
        raise SystemExit("oracle/_ref is not built: make -C oracle")
    out = {}
    names = []
    for name, lay, ain, is_s, om, nz, dz, zorig in hk_cases():
        names.append(name)
        for k in ("top", "vp", "vs", "rho", "vpd", "vsd"):
            out["%s__%s" % (name, k)] = np.asarray(lay[k], np.float64)
        out[name + "__scalars"] = np.array([ain, is_s, om, nz, dz, zorig], np.float64)
        for prec in (32, 64):
            ierr, u, w = O.ref_haskgrn(lay, ain, is_s, om, nz, dz, zorig, prec)
            out["%s__u%d" % (name, prec)] = u
            out["%s__w%d" % (name, prec)] = w
            out["%s__ierr%d" % (name, prec)] = np.array(ierr, np.int32)
    out["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "haskgrn_cases.npz"), **out)

    # scattered-field source on an 8 x 7 x 7 grid
    nx, ny, nz, dx, freq = 8, 7, 7, 250.0, 0.3
    pmld, pmlr = 500.0, 1e-3
    rng = np.random.default_rng(2024)
    z = np.arange(nz) * dx
    vpb1 = 4500.0 + 150.0 * z / dx
    vpb = np.broadcast_to(vpb1[None, None, :], (ny, nx, nz)).copy()
    vsb = vpb / 1.75
    rhob = 310.0 * vpb ** 0.25
    pert = 1.0 + 0.05 * rng.standard_normal((ny, nx, nz))
    vp, vs, rho = vpb * pert, vsb * pert, rhob * (1.0 + 0.3 * (pert - 1.0))
    ub, vb, wb = [(rng.standard_normal((ny, nx, nz)) + 1j * rng.standard_normal((ny, nx, nz))) for _ in range(3)]
    srcfn = 0.7 - 0.2j
    res = dict(vp=vp, vs=vs, rho=rho, vpb=vpb, vsb=vsb, rhob=rhob, ub=ub, vb=vb, wb=wb,
               params=np.array([nx, ny, nz, dx, freq, pmld, pmlr, srcfn.real, srcfn.imag], np.float64))
    for prec in (32, 64):
        ft = np.float32 if prec == 32 else np.float64
        ierr, sf = O.ref_hetfldsrc(nx, ny, nz, dx, freq, *(a.astype(ft) for a in (vp, vs, rho, vpb, vsb, rhob)),
                                   ub, vb, wb, srcfn, pmld, pmlr, prec=prec)
        assert ierr == 0
        res["sfld%d" % prec] = sf
    np.savez_compressed(os.path.join(HERE, "hetfldsrc_g877.npz"), **res)

    # receivers and hydrophones on a 12 x 11 x 10 grid
    nx, ny, nz, dx, freq = 12, 11, 10, 100.0, 0.4
    rng = np.random.default_rng(31)
    vp = 4000.0 + 500.0 * rng.random((ny, nx, nz))
    vs = vp / 1.8
    rho = 2500.0 + 100.0 * rng.random((ny, nx, nz))
    n = 3 * nx * ny * nz
    sfld = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ub, vb, wb = [rng.standard_normal((ny, nx, nz)) + 1j * rng.standard_normal((ny, nx, nz)) for _ in range(3)]
    res = dict(vp=vp, vs=vs, rho=rho, sfld=sfld, ub=ub, vb=vb, wb=wb, params=np.array([nx, ny, nz, dx, freq]))
    sets = [("nearest", 0, 0.5), ("sinc2", 2, 0.5), ("sinc4", 4, 0.5), ("dirac1", 1, 1.0), ("dirac13", 2, 1.3)]
    for tag, ireg, spread in sets:
        xs = rng.uniform(0.0, (nx - 1) * dx, 6)
        ys = rng.uniform(0.0, (ny - 1) * dx, 6)
        zs = np.concatenate([[0.0, 30.0], rng.uniform(0.0, (nz - 1) * dx, 4)])
        fac = rng.standard_normal(6) + 1j * rng.standard_normal(6)
        hx = rng.uniform(dx, (nx - 2) * dx, 3)
        hy = rng.uniform(dx, (ny - 2) * dx, 3)
        hz = np.array([0.0, 150.0, 420.0])
        rec = dict(x=xs, y=ys, z=zs, spread=spread, ireg=ireg, fac=fac)
        hyd = dict(x=hx, y=hy, z=hz, spread=spread, ireg=ireg)
        res[tag + "__stations"] = np.stack([xs, ys, zs])
        res[tag + "__fac"] = fac
        res[tag + "__hyd"] = np.stack([hx, hy, hz])
        res[tag + "__set"] = np.array([ireg, spread])
        for prec in (32, 64):
            ierr, ue, ve, we, pe = O.ref_makeuh(nx, ny, nz, dx, freq, vp, vs, rho, sfld, ub, vb, wb, rec, hyd, prec=prec)
            assert ierr == 0
            res["%s__out%d" % (tag, prec)] = np.concatenate([ue, ve, we, pe])
    res["sets"] = np.array([t for t, _, _ in sets])
    np.savez_compressed(os.path.join(HERE, "makeuh_g121110.npz"), **res)


if __name__ == "__main__":
    main()
