"""CPU restatement of receiver extraction -- TEST INFRASTRUCTURE ONLY.

Follows MAKEUH (makeuh.f:1-172), SRCSETUP (srcsetup.f:1-214, extraction
branch lsetup = .false.) and SRCREG (srcreg.f) in double precision, loop by
loop:

- region weights (srcreg.f:58-141): a point receiver (spread <= 0.5) is a
  product of Kaiser-windowed sincs over a (2 ireg + 1)^3 region, the window
  parameter from the table of Hicks (2002, Geophysics 67, 156-166) for
  accuracy up to half the Nyquist wavenumber and I0 from the polynomial
  approximations of Abramowitz & Stegun 9.8.1-9.8.2; a spread receiver is a
  product of discrete Dirac pulses (Alterman & Aboudi 1970) over
  (INT(4 spread) + 1)^3 nodes normalised to unit sum;
- extraction (srcsetup.f:68-212): nearest node for ireg = 0; otherwise the
  weighted sum over the region, points beyond a free surface mirrored back
  into the grid with a change of sign, points beyond other boundaries
  dropped;
- MAKEUH: receivers u, v, w (times the receiver response when recin is
  set); hydrophone pressure (lambda + 2 mu / 3) div u at the nearest node
  with central differences of extractions at +-d, over 2 d (-i omega).

Parity: pinned against the compiled reference (oracle/_ref, MAKEUH through
the harness) in FP64 and FP32 (tests/test_rcv_oracle.py).
"""
import math

import numpy as np

# Hicks (2002) Kaiser window parameter b per sinc half-length r (kmax = pi/2)
KAISER_B = {1: 1.24, 2: 2.94, 3: 4.53, 4: 6.31, 5: 7.91, 6: 9.42, 7: 10.95, 8: 12.53, 9: 14.09}


def bessel_i0(x):
    """Modified Bessel I0 by Abramowitz & Stegun 9.8.1 (|x| < 3.75) and 9.8.2."""
    ax = abs(x)
    if ax < 3.75:
        t = (x / 3.75) ** 2
        return 1.0 + t * (3.5156229 + t * (3.0899424 + t * (1.2067492 + t * (0.2659732 + t * (0.0360768
                                                                                             + t * 0.0045813)))))
    t = 3.75 / ax
    p = 0.39894228 + t * (0.01328592 + t * (0.00225319 + t * (-0.00157565 + t * (0.00916281 + t * (
        -0.02057706 + t * (0.02635537 + t * (-0.01647633 + t * 0.00392377)))))))
    return (math.exp(ax) / math.sqrt(ax)) * p


def kaiser_sinc(r, x):
    s = 1.0 if x <= 1e-10 else math.sin(math.pi * x) / (math.pi * x)
    b = KAISER_B.get(r, 14.18)
    lx = x / float(r)
    taper = bessel_i0(b * (1.0 - lx * lx) ** 0.5) / bessel_i0(b) if lx < 1 else 0.0
    return s * taper


def dirac(h, x):
    if x < h:
        return (1.0 - x * x / (2.0 * h * h)) / (2.0 * h)
    if x < 2.0 * h:
        return (1.0 + (x * x - 4.0 * x * h) / (4.0 * h * h)) / h
    return 0.0


def region_weights(ireg, spread, dx, dy, dz, xoff, yoff, zoff):
    """fs[iz, ix, iy] over the region (0-based offsets 0..m-1)."""
    if spread <= 0.5:
        m = 2 * ireg + 1
        c = ireg + 1
        fs = np.zeros((m, m, m))
        for iy in range(1, m + 1):
            wy = kaiser_sinc(ireg, abs(iy - (c + yoff / dy)))
            for ix in range(1, m + 1):
                wx = kaiser_sinc(ireg, abs(ix - (c + xoff / dx)))
                for iz in range(1, m + 1):
                    fs[iz - 1, ix - 1, iy - 1] = kaiser_sinc(ireg, abs(iz - (c + zoff / dz))) * wy * wx
        return fs
    m = int(4 * spread) + 1
    c = int(2 * spread) + 1
    fs = np.zeros((m, m, m))
    for iy in range(1, m + 1):
        for ix in range(1, m + 1):
            for iz in range(1, m + 1):
                fs[iz - 1, ix - 1, iy - 1] = (dirac(spread, abs(iz - (c + zoff / dz))) *
                                              dirac(spread, abs(iy - (c + yoff / dy))) *
                                              dirac(spread, abs(ix - (c + xoff / dx))))
    return fs / fs.sum()


def extract(field, x, y, z, spread, ireg, dx, dy, dz, freesurf):
    """Value of `field` (ny, nx, nz) (z fastest) at (x, y, z); freesurf[6]
    as MODEL_MODULE freesurf (1 top, 2 +x, 3 bottom, 4 -x, 5 +y, 6 -y)."""
    ny, nx, nz = field.shape
    F = lambda iz, ix, iy: field[iy - 1, ix - 1, iz - 1]  # noqa: E731
    if ireg == 0:
        return F(int(0.5 + z / dz) + 1, int(0.5 + x / dx) + 1, int(0.5 + y / dy) + 1)
    ixs, iys, izs = int(x / dx) + 1, int(y / dy) + 1, int(z / dz) + 1
    xoff, yoff, zoff = x - (ixs - 1) * dx, y - (iys - 1) * dy, z - (izs - 1) * dz
    h = ireg if spread <= 0.5 else int(2 * spread)
    lo = {"x": ixs - h, "y": iys - h, "z": izs - h}
    hi = {"x": ixs + h, "y": iys + h, "z": izs + h}
    fs = region_weights(ireg, spread, dx, dy, dz, xoff, yoff, zoff)
    zmin = lo["z"] if freesurf[0] else max(lo["z"], 1)
    zmax = hi["z"] if freesurf[2] else min(hi["z"], nz)
    xmin = lo["x"] if freesurf[3] else max(lo["x"], 1)
    xmax = hi["x"] if freesurf[1] else min(hi["x"], nx)
    ymin = lo["y"] if freesurf[5] else max(lo["y"], 1)
    ymax = hi["y"] if freesurf[4] else min(hi["y"], ny)

    def mirror(i, n):
        return 2 - i if i < 1 else (2 * n - i if i > n else i)

    acc = 0.0 + 0.0j
    for iz in range(zmin, zmax + 1):
        izf = mirror(iz, nz)
        for ix in range(xmin, xmax + 1):
            ixf = mirror(ix, nx)
            for iy in range(ymin, ymax + 1):
                iyf = mirror(iy, ny)
                w = fs[iz - lo["z"], ix - lo["x"], iy - lo["y"]]
                v = F(izf, ixf, iyf)
                acc = acc + w * v if (iz == izf and ix == ixf and iy == iyf) else acc - w * v
    return acc


def makeuh(u, v, w, receivers, hydrophones, lam, mu, dx, omega, freesurf):
    """Receiver displacements and hydrophone pressures from the total field.

    u, v, w: (ny, nx, nz) field components; receivers: dict x, y, z,
    spread, ireg, fac (None or per-receiver response); hydrophones: dict x,
    y, z, spread, ireg; lam, mu: (ny, nx, nz) complex moduli.  Returns
    (uest, vest, west, utest)."""
    ue, ve, we = [], [], []
    for k in range(len(receivers["x"])):
        args = (receivers["x"][k], receivers["y"][k], receivers["z"][k], receivers["spread"], receivers["ireg"],
                dx, dx, dx, freesurf)
        f = 1.0 if receivers.get("fac") is None else receivers["fac"][k]
        ue.append(extract(u, *args) * f)
        ve.append(extract(v, *args) * f)
        we.append(extract(w, *args) * f)
    ny, nx, nz = u.shape
    cfact = complex(0.0, -1.0) * omega
    pe = []
    for k in range(len(hydrophones["x"])):
        x, y, z = hydrophones["x"][k], hydrophones["y"][k], hydrophones["z"][k]
        sp, ir = hydrophones["spread"], hydrophones["ireg"]
        ix = min(max(int(1.5 + x / dx), 1), nx)
        iy = min(max(int(1.5 + y / dx), 1), ny)
        iz = min(max(int(1.5 + z / dx), 1), nz)
        ex = lambda fld, a, b, c: extract(fld, a, b, c, sp, ir, dx, dx, dx, freesurf)  # noqa: E731
        dudx = -(ex(u, x + dx, y, z) - ex(u, x - dx, y, z)) / (2.0 * dx * cfact)
        dvdy = -(ex(v, x, y + dx, z) - ex(v, x, y - dx, z)) / (2.0 * dx * cfact)
        dwdz = -(ex(w, x, y, z + dx) - ex(w, x, y, z - dx)) / (2.0 * dx * cfact)
        pe.append((lam[iy - 1, ix - 1, iz - 1] + 2.0 * mu[iy - 1, ix - 1, iz - 1] / 3.0) * (dudx + dvdy + dwdz))
    return np.array(ue), np.array(ve), np.array(we), np.array(pe)
