/* CPU oracle for the wtomo3d assembly path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference's graph, CSR and stencil assembly,
 * used by tests/, __graft_entry__.smoke() and bench.py's CPU baseline as the
 * checker.  The product (wtomo3d_amd) never links or calls it.
 *
 *   ora_graph_size / ora_graph   srgraphs.f:626-734 (SREL3D_GETSIZE, SREL3D)
 *                                + sparse.f:13-49 (SPARSE_ADJ2CRS), 1-based
 *   ora_assemble_{f32,f64}       asmble.f90:96-601 with setcoefs3d.f:79-233,
 *                                calcwts.f:28-71, fdfd3d.f:50-115,
 *                                pfdfd3d.f:133-576, pml3d.f:167-800,
 *                                fixed3d.f:41-343
 *   ora_pml_params               readini.f90:81-84
 *   ora_lame_{f32,f64}           modelio.f90:78-92
 *
 * Parity pins: the pattern against the compiled reference (oracle/_ref) and
 * closed forms; values against the compiled reference in FP32 (as shipped)
 * and FP64 (-fdefault-real-8 -fdefault-double-8), plus invariants in tests/.
 * The traction-free top (tpfree.f, usemin = false) is not restated: the
 * assembler returns 2 for it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

#define R float
#define SFX f32
#define FABS fabsf
#define SQRT sqrtf
#define HYPOT hypotf
#include "asm_oracle_body.h"
#undef R
#undef SFX
#undef FABS
#undef SQRT
#undef HYPOT

#define R double
#define SFX f64
#define FABS fabs
#define SQRT sqrt
#define HYPOT hypot
#include "asm_oracle_body.h"
#undef R
#undef SFX

/* modelio.f90:78-92: mu = vs^2 rho, lambda = vp^2 rho - 2 mu, in the
 * reference's complex arithmetic (imaginary parts are exactly zero). */
#define LAME(R, SFX)                                                          \
void CAT(ora_lame_, SFX)(long n, const R *vp, const R *vs, const R *rho,     \
                         R *lam, R *mu)                                       \
{                                                                             \
    for (long g = 0; g < n; ++g) {                                            \
        R m = vs[g] * vs[g]; m = m * rho[g];                                   \
        R l = vp[g] * vp[g]; l = l * rho[g]; l = l - (R)2 * m;                 \
        mu[2 * g] = m; mu[2 * g + 1] = 0;                                      \
        lam[2 * g] = l; lam[2 * g + 1] = 0;                                    \
    }                                                                         \
}
LAME(float, f32)
LAME(double, f64)

/* readini.f90:81-84 in single (f32) or double (f64) precision */
void ora_pml_params_f32(float dx, float pmld, float pmlr, int *ipml, float *pmlf)
{
    *ipml = (int)(pmld / dx) + 1;
    *pmlf = 3.0f * logf(1.0f / pmlr) / (2.0f * pmld * pmld * pmld);
}
void ora_pml_params_f64(double dx, double pmld, double pmlr, int *ipml, double *pmlf)
{
    *ipml = (int)(pmld / dx) + 1;
    *pmlf = 3.0 * log(1.0 / pmlr) / (2.0 * pmld * pmld * pmld);
}

/* srgraphs.f:626-636: number of unknowns and of nonzeros (2*nedge + n) */
void ora_graph_size(int nx, int ny, int nz, long long *n, long long *nnz)
{
    long long X = nx, Y = ny, Z = nz;
    long long nedge = 120 * X * Y * Z - 81 * (X * Y + X * Z + Y * Z) + 54 * (X + Z + Y) - 36;
    *n = 3 * X * Y * Z;
    *nnz = 2 * nedge + *n;
}

/* srgraphs.f:638-734 (adjacency without the diagonal, neighbours y > x > z >
 * component) followed by sparse.f:13-49 (insert the diagonal before the
 * first larger column; the last row's diagonal is appended).  1-based. */
int ora_graph(int nx, int ny, int nz, int32_t *irptr, int32_t *jcptr)
{
    long long n, nnz;
    ora_graph_size(nx, ny, nz, &n, &nnz);
    long long pos = 0, row = 0;
    for (int iy = 1; iy <= ny; ++iy) {
        int ym = iy > 1 ? iy - 1 : 1, yp = iy < ny ? iy + 1 : ny;
        for (int ix = 1; ix <= nx; ++ix) {
            int xm = ix > 1 ? ix - 1 : 1, xp = ix < nx ? ix + 1 : nx;
            for (int iz = 1; iz <= nz; ++iz) {
                int zm = iz > 1 ? iz - 1 : 1, zp = iz < nz ? iz + 1 : nz;
                long long ng = iz + (long long)(ix - 1) * nz + (long long)(iy - 1) * nx * nz;
                for (int r = 1; r <= 3; ++r) {
                    long long i = 3 * ng - 3 + r;          /* this row, 1-based */
                    int diag = 0;
                    irptr[row++] = (int32_t)(pos + 1);
                    for (int j = ym; j <= yp; ++j)
                        for (int ii = xm; ii <= xp; ++ii)
                            for (int k = zm; k <= zp; ++k) {
                                long long g = k + (long long)(ii - 1) * nz + (long long)(j - 1) * nx * nz;
                                for (int c = 1; c <= 3; ++c) {
                                    long long col = 3 * g - 3 + c;
                                    if (col == i) continue;
                                    if (!diag && col > i) { jcptr[pos++] = (int32_t)i; diag = 1; }
                                    jcptr[pos++] = (int32_t)col;
                                }
                            }
                    if (!diag) jcptr[pos++] = (int32_t)i;   /* only the last row */
                }
            }
        }
    }
    irptr[row] = (int32_t)(pos + 1);
    return pos == nnz ? 0 : 1;
}
