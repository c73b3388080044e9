"""CPU restatement of the scattered-field source -- TEST INFRASTRUCTURE ONLY.

HETFLDSRC (hetfldsrc.f90:17-1039) forms, for every node, the 27-point
stencil of the background model minus that of the current model (entry by
entry: hetfldsrc.f90:415-426 saves the current model's coefficients, then
subtracts them from the background's) applied to the background fields,
times srcfn (hetfldsrc.f90:1021-1024).  The stencil entries are ASMBLE's
(the same SETCOEFS3D/FDFD3D calls, asmble.f90:96-601), so with both
operators in CSR form the source is

    sfld = srcfn * (A_bg - A) u_bg

over the CSR pattern (neighbours outside the grid are absent in both).
Node numbering is the unknown order; the reference's own numbering
(hetfldsrc.f90:1020 uses nx*ny where the unknown order has nx*nz) agrees
only when ny = nz.
Parity: pinned against the compiled reference (oracle/_ref) in FP64 and
FP32 on an ny = nz grid, and by the invariants A_bg = A => 0 and
linearity in srcfn (tests/test_rhs_oracle.py).
"""
import numpy as np
import scipy.sparse as sp


def scattered_source(irptr, jcptr, a, abg, u, srcfn=1.0):
    """irptr/jcptr 1-based CSR, a/abg the two operators' values, u (n,) or
    (n, nrhs); returns srcfn * (abg - a) u."""
    n = irptr.size - 1
    D = sp.csr_matrix((np.asarray(abg) - np.asarray(a), np.asarray(jcptr) - 1, np.asarray(irptr) - 1),
                      shape=(n, n))
    return np.asarray(srcfn) * (D @ u)


def nodes_to_unknowns(ub, vb, wb):
    """Per-component grids (ny, nx, nz) (z fastest) -> unknown vector
    (u, v, w per node)."""
    return np.stack([np.ravel(ub), np.ravel(vb), np.ravel(wb)], axis=1).ravel()


def planewave_expand(nx, ny, nz, dx, U, W, kh, phi):
    """The 3-D background field of plane waves from their depth profiles (the
    expansion the reference's missing planewvexp.f90 would do, defined here
    from its interface, interface.f90:43-47):

        u_j = (U_j(z) cos phi_j, U_j(z) sin phi_j, W_j(z))
              * exp(-i kh_j (x cos phi_j + y sin phi_j)),

    x = ix dx, y = iy dx (0-based nodes), phi_j the azimuth of horizontal
    propagation from +x towards +y, U along it, W vertical (positive down),
    the e^{+i omega t} convention of the profiles.  U, W: (nrhs, nz) complex;
    kh, phi: (nrhs,).  Returns row-major (n, nrhs) complex128 in the unknown
    order (u, v, w per node; z fastest, then x, then y)."""
    U = np.atleast_2d(np.asarray(U, np.complex128))
    W = np.atleast_2d(np.asarray(W, np.complex128))
    kh = np.atleast_1d(np.asarray(kh, np.float64))
    phi = np.atleast_1d(np.asarray(phi, np.float64))
    c, s = np.cos(phi), np.sin(phi)
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    # phase[iy, ix, j]
    ph = np.exp(-1j * kh[None, None, :] * (x[None, :, None] * c[None, None, :] + y[:, None, None] * s[None, None, :]))
    out = np.empty((ny, nx, nz, 3, kh.size), np.complex128)
    out[..., 0, :] = ph[:, :, None, :] * (U.T * c)[None, None, :, :]
    out[..., 1, :] = ph[:, :, None, :] * (U.T * s)[None, None, :, :]
    out[..., 2, :] = ph[:, :, None, :] * W.T[None, None, :, :]
    return out.reshape(-1, kh.size)
