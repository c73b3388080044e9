"""Synthetic elastic media and sources for the benchmark configurations.

Velocity/density layouts follow the reference's model arrays: (nz, nx, ny)
Fortran order, i.e. numpy arrays of shape (ny, nx, nz) in C order (z
fastest), as MODELIO_READ_MODEL stores them (modelio.f90:98-113).
"""
import numpy as np


def layered(nx, ny, nz, dx, seed, nlayers=6):
    """1-D layered model: seeded random layer thicknesses, Vp uniform in
    4-8 km/s increasing with depth, Vs = Vp/1.73, Gardner density
    rho = 310 Vp^0.25 (kg/m^3, Vp in m/s).  Returns (vp, vs, rho) 1-D in z."""
    rng = np.random.default_rng(seed)
    z = np.arange(nz) * dx
    tops = np.sort(rng.uniform(0.0, nz * dx, nlayers - 1))
    vps = np.sort(rng.uniform(4000.0, 8000.0, nlayers))
    vp = vps[np.searchsorted(tops, z, side="right")]
    return vp, vp / 1.73, 310.0 * vp ** 0.25


def layer_stack(nx, ny, nz, dx, seed, nlayers=6):
    """The layers of ``layered`` (same seed, same draws) as a 1-D stack for
    the Thomson-Haskell profiles: dict of per-layer arrays ``top`` (depth of
    each layer's top, the first 0), ``vp``, ``vs``, ``rho``; the last layer
    is the half-space.  The discrete speeds ``vpd``/``vsd`` are the same."""
    rng = np.random.default_rng(seed)
    tops = np.sort(rng.uniform(0.0, nz * dx, nlayers - 1))
    vps = np.sort(rng.uniform(4000.0, 8000.0, nlayers))
    top = np.concatenate([[0.0], tops])
    vs = vps / 1.73
    return dict(top=top, vp=vps, vs=vs, rho=310.0 * vps ** 0.25, vpd=vps.copy(), vsd=vs.copy())


def background_1d(vp, vs, rho, dx, max_layers=64):
    """Layered background of a 3-D model (arrays (ny, nx, nz), z fastest):
    the lateral mean of vp, vs and rho at each depth, with runs of equal
    consecutive depths merged into layers.  When that leaves more than
    ``max_layers`` layers (the profile generator's limit), the depths are cut
    into ``max_layers`` runs of near-equal length and each run takes its mean.
    Layer tops lie half a cell above a layer's first node (the first at 0), so
    no node sits on an interface.  A one-layer background gets a second,
    identical layer below the grid (the profiles need a layer above the
    half-space).  Returns (vp1, vs1, rho1) per depth -- piecewise constant,
    the values of the layer holding each node -- and the layer stack in the
    form of ``layer_stack``."""
    def mean(v, axis=None):
        # a run of equal values keeps its value exactly
        v = np.asarray(v, np.float64)
        return np.where(v.max(axis=axis) == v.min(axis=axis), v.max(axis=axis), v.mean(axis=axis))

    m = [mean(np.reshape(a, (-1, np.shape(a)[-1])), axis=0) for a in (vp, vs, rho)]
    nz = m[0].size
    starts = [0] + [iz for iz in range(1, nz) if any(a[iz] != a[iz - 1] for a in m)]
    if len(starts) > max_layers:
        starts = sorted(set(int(v) for v in np.linspace(0, nz, max_layers, endpoint=False)))
    bounds = starts + [nz]
    per_layer = [np.array([mean(a[bounds[k]:bounds[k + 1]]) for k in range(len(starts))]) for a in m]
    per_depth = [np.repeat(v, np.diff(bounds)) for v in per_layer]
    top = np.array([0.0] + [(s - 0.5) * dx for s in starts[1:]])
    if len(starts) == 1:
        per_layer = [np.repeat(v, 2) for v in per_layer]
        top = np.array([0.0, nz * dx])
    vpl, vsl, rhol = per_layer
    layers = dict(top=top, vp=vpl, vs=vsl, rho=rhol, vpd=vpl.copy(), vsd=vsl.copy())
    return per_depth[0], per_depth[1], per_depth[2], layers


def gaussian_field(shape, dx, corr_len, std, seed):
    """Zero-mean Gaussian random field with Gaussian covariance of the given
    correlation length, scaled to the requested standard deviation
    (spectral filtering of seeded white noise)."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n, d=dx) * 2 * np.pi for n in shape]
    K2 = k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2
    amp = np.exp(-K2 * corr_len ** 2 / 8.0)  # sqrt of the Gaussian spectrum
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * amp))
    f -= f.mean()
    return f * (std / f.std())


def heterogeneous(nx, ny, nz, dx, seed, std=0.05, corr_len=2000.0):
    """Layered background plus a Gaussian random field on Vp and Vs (std,
    correlation length) with the density perturbed by 0.3 times the Vp
    perturbation.  Returns (vp, vs, rho) of shape (ny, nx, nz)."""
    vp1, vs1, rho1 = layered(nx, ny, nz, dx, seed)
    d = gaussian_field((ny, nx, nz), dx, corr_len, std, seed + 7)
    vp = vp1[None, None, :] * (1.0 + d)
    vs = vs1[None, None, :] * (1.0 + d)
    rho = rho1[None, None, :] * (1.0 + 0.3 * d)
    return vp, vs, rho


def lame(vp, vs, rho):
    """(lambda, mu) complex128 and rho float64, flattened z-fastest, formed
    as MODELIO_READ_MODEL does (modelio.f90:78-92)."""
    vp = np.asarray(vp, np.float64).ravel()
    vs = np.asarray(vs, np.float64).ravel()
    rho = np.asarray(rho, np.float64).ravel()
    mu = (vs * vs) * rho
    lam = (vp * vp) * rho - 2.0 * mu
    return lam.astype(np.complex128), mu.astype(np.complex128), rho


def plane_wave_fields(nx, ny, nz, dx, omega, vp, nrhs, seed, inc=(10.0, 35.0)):
    """Displacement fields of homogeneous P plane waves arriving from below
    (incidence uniform in `inc` degrees, azimuth uniform in 0-360), one per
    column: u(r) = d exp(-i omega d.r / vp), d the propagation direction
    (z positive down).  Row-major (n, nrhs) complex128 in the reference's
    unknown order (u, v, w per node, z fastest, then x, then y)."""
    rng = np.random.default_rng(seed)
    th = np.deg2rad(rng.uniform(inc[0], inc[1], nrhs))
    ph = np.deg2rad(rng.uniform(0.0, 360.0, nrhs))
    d = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), -np.cos(th)], axis=1)  # (nrhs, 3)
    iy, ix, iz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    r = np.stack([ix.ravel() * dx, iy.ravel() * dx, iz.ravel() * dx], axis=1)  # nodes, z fastest
    phase = np.exp(-1j * (omega / vp) * (r @ d.T))  # (nodes, nrhs)
    U = phase[:, None, :] * d.T[None, :, :]        # (nodes, 3, nrhs)
    return np.ascontiguousarray(U.reshape(-1, nrhs))


def plane_wave_dirs(nrhs, seed, inc=(10.0, 35.0)):
    """Unit propagation directions (z down, arriving from below) of nrhs
    plane waves with incidence uniform in `inc` degrees and azimuth uniform
    in 0-360 degrees: array (nrhs, 3)."""
    rng = np.random.default_rng(seed)
    th = np.deg2rad(rng.uniform(inc[0], inc[1], nrhs))
    ph = np.deg2rad(rng.uniform(0.0, 360.0, nrhs))
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), -np.cos(th)], axis=1)


def attenuating(vp, vs, qpi, qsi, rho, freqbase, omega):
    """Complex (lambda, mu) of an attenuating model at angular frequency
    omega (DISPERSION, dispersion.f:1-131), formed on the GPU: vp, vs the
    speeds at `freqbase` Hz, qpi, qsi = 1/Qp, 1/Qs.  Returns complex128
    arrays flattened z fastest, with rho as float64."""
    import ctypes
    from ._lib import lib, check
    arrs = [np.ascontiguousarray(np.asarray(a, np.float64).ravel()) for a in (vp, vs, qpi, qsi, rho)]
    n = arrs[0].size
    lam = np.zeros(n, np.complex128)
    mu = np.zeros(n, np.complex128)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    check(lib().wt_dispersion(n, *[P(a) for a in arrs], float(freqbase), float(np.real(omega)), P(lam), P(mu), 0,
                              None), "wt_dispersion")
    return lam, mu, arrs[4]
