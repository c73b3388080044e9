"""Write the assembly golden fixtures from the compiled reference.

Runs only where oracle/_ref is built (the reference's own Fortran sources,
compiled by oracle/Makefile).  Each case stores its inputs (velocity, density,
parameters) and the reference outputs: the 1-based CSR pattern and the
complex64 (as shipped) and complex128 (-fdefault-real-8 -fdefault-double-8)
values.  The 24^3 case (the plain-method 24x24x24 configuration) keeps only
SHA-256 digests of the pattern and the values of a row sample.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

# (name, nx, ny, nz, dx, freq, pmld, pmlr, fst, tau, seed, usemin); usemin = 0
# is the traction-free top (TPFREE, fdfd3d.f:73-85)
CASES = [
    ("g667_fs", 6, 6, 7, 250.0, 0.3, 250.0, 1e-3, 1, 999.999, 11, 1),
    ("g667_fixed", 6, 6, 7, 250.0, 0.3, 250.0, 1e-3, 0, 999.999, 12, 1),
    ("g667_tau", 6, 6, 7, 250.0, 0.25, 250.0, 1e-3, 1, 3.0, 13, 1),
    ("g6610_thinz", 6, 6, 10, 400.0, 0.15, 4000.0, 1e-3, 1, 999.999, 14, 1),
    ("g987_tpfree", 9, 8, 7, 250.0, 0.3, 500.0, 1e-3, 1, 999.999, 15, 0),
    ("g667_tpfree_tau", 6, 6, 7, 250.0, 0.25, 250.0, 1e-3, 1, 3.0, 16, 0),
]


def model(nx, ny, nz, seed, prec):
    ft = np.float32 if prec == 32 else np.float64
    rng = np.random.default_rng(seed)
    shape = (ny, nx, nz)  # C order, z fastest == Fortran (nz, nx, ny)
    vp = 4000.0 + 3000.0 * rng.random(shape)
    vs = vp / (1.6 + 0.3 * rng.random(shape))
    rho = 2000.0 + 1000.0 * rng.random(shape)
    return vp.astype(ft), vs.astype(ft), rho.astype(ft)


def main():
    if not O.have_ref():
        raise SystemExit("oracle/_ref is not built: make -C oracle")
    meta = {}
    for name, nx, ny, nz, dx, freq, pmld, pmlr, fst, tau, seed, usemin in CASES:
        out = {}
        for prec in (32, 64):
            vp, vs, rho = model(nx, ny, nz, seed, prec)
            ierr, ir, jc, a = O.ref_assemble(nx, ny, nz, dx, freq, vp, vs, rho, pmld, pmlr,
                                             fst=bool(fst), usemin=bool(usemin), prec=prec, tau=tau)
            assert ierr == 0
            out["vals%d" % prec] = a
            if prec == 64:
                out.update(vp=vp, vs=vs, rho=rho, irptr=ir, jcptr=jc)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        meta[name] = dict(nx=nx, ny=ny, nz=nz, dx=dx, freq=freq, pmld=pmld, pmlr=pmlr, fst=fst,
                          tau=tau, seed=seed, usemin=usemin)
    # 24^3 plain-method configuration: digests + sampled rows
    nx = ny = nz = 24
    ir, jc = O.ref_graph(nx, ny, nz)
    meta["c1_pattern_sha256"] = dict(irptr=hashlib.sha256(ir.tobytes()).hexdigest(),
                                     jcptr=hashlib.sha256(jc.tobytes()).hexdigest(),
                                     n=int(ir.size - 1), nnz=int(jc.size))
    rng = np.random.default_rng(1001)
    rows = np.sort(rng.choice(3 * nx * ny * nz, 2000, replace=False))
    samp = {"rows": rows}
    for prec in (32, 64):
        ft = np.float32 if prec == 32 else np.float64
        vp = np.full((ny, nx, nz), 6000.0, ft)
        vs = (vp / ft(np.sqrt(3.0))).astype(ft)
        rho = np.full_like(vp, 2700.0)
        ierr, ir2, jc2, a = O.ref_assemble(nx, ny, nz, 500.0, 0.2, vp, vs, rho, 5000.0, 1e-3,
                                           fst=True, prec=prec)
        assert ierr == 0
        idx = np.concatenate([np.arange(ir2[r] - 1, ir2[r + 1] - 1) for r in rows])
        samp["vals%d" % prec] = a[idx]
    np.savez_compressed(os.path.join(HERE, "c1_rows.npz"), **samp)
    meta["c1_rows"] = dict(nx=24, ny=24, nz=24, dx=500.0, freq=0.2, pmld=5000.0, pmlr=1e-3, fst=1,
                           vp=6000.0, vs="6000/sqrt(3)", rho=2700.0)
    with open(os.path.join(HERE, "assembly_cases.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
