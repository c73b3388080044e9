/*
 * wtomo3d MI355X library: C ABI.
 *
 * Complex arrays are interleaved (re, im) pairs: float[2] per complex64,
 * double[2] per complex128 -- the memory layout of Fortran COMPLEX and C99
 * _Complex.  Model arrays are (nz, nx, ny) with z fastest, the layout of the
 * reference's MODEL_MODULE arrays (modules.F90:12-14).  Unknowns are ordered
 * 3*(iz + nz*(ix-1) + nx*nz*(iy-1)) - {2,1,0} for (u, v, w) as in
 * srgraphs.f:659-665.  Every entry point returns 0 on success or a WT_ERR_*
 * code; nothing here allocates on behalf of the caller unless it says so.
 */
#ifndef WTOMO3D_H
#define WTOMO3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid / medium / boundary parameters of one assembly.  The reference keeps
 * them in module globals: MODEL_MODULE (nx, ny, nz, dx, omega, freesurf),
 * PML_MODULE (pmld, pmlf, ipml) and INIT_MODULE (usemin), modules.F90:1-76. */
typedef struct wt_grid_params {
    int32_t nx, ny, nz;
    int32_t ipml;      /* INT(pmld/dx) + 1, readini.f90:83 */
    int32_t fst;       /* freesurf(1): free surface at iz = 1 */
    int32_t usemin;    /* Min-cell free surface; 0 (traction-free) is unsupported */
    double dx;         /* grid spacing [m]; dx = dy = dz (readini.f90:55-59) */
    double pmld;       /* PML width [m] */
    double pmlf;       /* 3 ln(1/pmlr) / (2 pmld^3), readini.f90:84 */
    double omega_re;   /* angular frequency 2 pi f (elwave.f90:151) */
    double omega_im;   /* + 1/tau when the anti-alias damping is on (elwave.f90:164) */
} wt_grid_params;

#define WT_PREC_F32 32
#define WT_PREC_F64 64

/* ---------------- CSR pattern and assembly ---------------- */

/* Number of nonzeros, 9(3nx-2)(3ny-2)(3nz-2) = 2*nedge + n
 * (SREL3D_GETSIZE, srgraphs.f:626-636). */
int64_t wt_graph_nnz(int nx, int ny, int nz);

/* CSR pattern with the given index base (1 reproduces the reference):
 * replaces SREL3D + SPARSE_ADJ2CRS (srgraphs.f:638-734, sparse.f:13-49;
 * called at elwave.f90:92-124).  Host arrays rowptr[3nxnynz+1],
 * colidx[nnz]; built on the GPU. */
int wt_graph(int nx, int ny, int nz, int base, int32_t *rowptr, int32_t *colidx);

/* Same, into device arrays, on a HIP stream (NULL = default stream). */
int wt_graph_dev(int nx, int ny, int nz, int base, int32_t *d_rowptr, int32_t *d_colidx,
                 void *stream);

/* CSR values of the elastic stencil + PML operator: replaces the numerics
 * of ASMBLE (asmble.f90:1-636).  prec = WT_PREC_F32 (complex64 in/out, the
 * reference's precision) or WT_PREC_F64 (complex128).  Host arrays:
 * lam, mu: complex[nxnynz]; rho: real[nxnynz]; vals: complex[nnz]. */
int wt_assemble(const wt_grid_params *p, int prec, const void *lam, const void *mu,
                const void *rho, void *vals);

/* Same on device arrays and a HIP stream; *d_err (device int, may be NULL)
 * receives bit 1 for a negative PML depth (pml3d.f:219-223). */
int wt_assemble_dev(const wt_grid_params *p, int prec, const void *d_lam, const void *d_mu,
                    const void *d_rho, void *d_vals, int *d_err, void *stream);

/* Drop-in for the reference's ASMBLE argument list (asmble.f90:1-2,
 * interface.f90:4-12) plus the module state it reads, passed explicitly by
 * the Fortran shim wtomo3d_amd/fortran/asmble.f90.  Checks the caller's
 * row pointers and nonzero count like asmble.f90:556-564,619-626 and sets
 * *ierr != 0 on a mismatch.  The device copy of the matrix stays resident
 * for strumpack_serial_finter. */
void wt_asmble(int lunlog, int nzero, int nx, int ny, int nz, const int32_t *irptr,
               const int32_t *jcptr, void *a, int *ierr, const void *da, const void *mu,
               const float *rho, float omega_re, float omega_im, float dx, float pmld,
               float pmlf, int ipml, int fst, int usemin);

/* ---------------- device memory helpers ---------------- */
int wt_dev_alloc(size_t bytes, void **d_ptr);
int wt_dev_free(void *d_ptr);
int wt_h2d(void *d_dst, const void *src, size_t bytes);
int wt_d2h(void *dst, const void *d_src, size_t bytes);
int wt_dev_sync(void);
/* Number of visible GPUs (0 without one). */
int wt_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* WTOMO3D_H */
