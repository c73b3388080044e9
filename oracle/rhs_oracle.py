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
