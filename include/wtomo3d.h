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

/* Return codes. */
#define WT_OK 0
#define WT_ERR_ARG 1          /* invalid argument or size */
#define WT_ERR_HIP 2          /* a HIP runtime call failed */
#define WT_ERR_STATE 3        /* call out of order (no model, values, factor ...) */
#define WT_ERR_UNSUPPORTED 4  /* a configuration the library does not build */
#define WT_ERR_NOMEM 5        /* device allocation failed */
#define WT_ERR_PATTERN 6      /* CSR pattern is not the 27-point elastic grid's */
#define WT_ERR_PIVOT 7        /* zero pivot, or a non-finite residual */
#define WT_ERR_NOCONV 8       /* refinement stopped above tol (max_iter reached or the
                               * residual stopped falling); the solution and the
                               * statistics are still written */

/* Grid / medium / boundary parameters of one assembly.  The reference keeps
 * them in module globals: MODEL_MODULE (nx, ny, nz, dx, omega, freesurf),
 * PML_MODULE (pmld, pmlf, ipml) and INIT_MODULE (usemin), modules.F90:1-76. */
typedef struct wt_grid_params {
    int32_t nx, ny, nz;
    int32_t ipml;      /* INT(pmld/dx) + 1, readini.f90:83 */
    int32_t fst;       /* freesurf(1): free surface at iz = 1 */
    int32_t usemin;    /* 1: Min-cell free surface; 0: traction-free top rows (TPFREE,
                        * tpfree.f:55-420, at iz = 1 when fst is set) */
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

/* ---------------- multifrontal solver ---------------- */

/* Opaque solver handle: one grid (nx, ny, nz), its nested-dissection
 * elimination tree and the factor / solve buffers on the GPU. */
typedef struct wt_solver wt_solver;

typedef struct wt_solve_stats {
    int32_t iterations;          /* solves with the factor (1 = no refinement) */
    int32_t nrhs;
    double rel_residual;         /* max_j ||b_j - A x_j|| / ||b_j|| (fp64) */
    double rel_residual_first;   /* the same after the first solve */
    double abs_residual;         /* max_j ||b_j - A x_j|| */
    double ms;                   /* wall time of the call */
    int32_t converged;           /* 1: rel_residual <= tol at exit (always 1 for tol <= 0) */
    int32_t stalled;             /* 1: stopped because the residual did not fall */
} wt_solve_stats;

/* Ordering + symbolic analysis (the work of STRUMPACK_reorder in job 1,
 * strumpack_finter.c:24-66): geometric nested dissection of the grid down
 * to boxes of at most leaf_nodes nodes (<= 0: 64), front sizes, update sets,
 * assembly maps and arena offsets.  Builds the CSR pattern on the GPU. */
int wt_solver_create(int nx, int ny, int nz, int leaf_nodes, wt_solver **out);
int wt_solver_destroy(wt_solver *h);

/* info[8] = {unknowns, nonzeros, fronts, tree levels, max pivots per front,
 * max update size, factor entries, contribution-block arena entries};
 * *flops = sum over fronts of 8 (p^3/3 + p^2 u + p u^2). */
int wt_solver_info(const wt_solver *h, int64_t *info, double *flops);

/* The same statistics from the host symbolic analysis alone (no GPU). */
int wt_symbolic_info(int nx, int ny, int nz, int leaf_nodes, int64_t *info, double *flops);

/* HBM bytes a solver for this grid will need (host-only symbolic
 * analysis; used to admit requests before anything is allocated). */
int64_t wt_estimate_bytes(int nx, int ny, int nz, int leaf_nodes, int factor_prec, int nrhs);

/* HBM bytes the solver needs for a factorization in factor_prec and a
 * block solve of nrhs columns (the admission estimate). */
int64_t wt_solver_bytes(const wt_solver *h, int factor_prec, int nrhs);

/* lambda, mu: complex128[nxnynz]; rho: double[nxnynz] (host or device). */
int wt_solver_set_model(wt_solver *h, const void *lam, const void *mu, const double *rho, int on_device);

/* Complex moduli of an attenuating model at one frequency; replaces
 * DISPERSION (dispersion.f:1-131; module inputs MODEL_MODULE vpr, vsr, qp,
 * qs, rho, omega, freqbase):
 *   c = c_base max(1 + ln(omega_re / (2 pi freqbase)) / (pi Q), 0.1),
 *   c~ = c (1 - i / (2 Q)),  mu = vs~^2 rho,  lambda = vp~^2 rho - 2 mu,
 * vp, vs the speeds at freqbase, qpi, qsi = 1/Qp, 1/Qs (the reference's qp,
 * qs arrays), rho; all double [nodes].  When omega_re or freqbase is not
 * positive the moduli are those of the base speeds (the reference leaves
 * them unchanged).  lam, mu: complex128 [nodes].  on_device: every array is
 * a device pointer; otherwise host arrays.  Runs on the given stream. */
int wt_dispersion(int64_t nodes, const double *vp, const double *vs, const double *qpi, const double *qsi,
                  const double *rho, double freqbase, double omega_re, void *lam, void *mu, int on_device,
                  void *stream);

/* wt_solver_set_model with the moduli of wt_dispersion, formed on the
 * device from host vp, vs, 1/Qp, 1/Qs, rho. */
int wt_solver_set_model_q(wt_solver *h, const double *vp, const double *vs, const double *qpi,
                          const double *qsi, const double *rho, double freqbase, double omega_re);

/* FP64 assembly of the operator at p->omega into the solver's CSR values. */
int wt_solver_assemble(wt_solver *h, const wt_grid_params *p);

/* CSR values in the library's pattern order (wt_graph), complex64
 * (prec 32, widened) or complex128 (prec 64); host or device. */
int wt_solver_set_values(wt_solver *h, const void *vals, int prec, int on_device);
int wt_solver_get_values(wt_solver *h, void *vals);

/* Multifrontal LU of the current values in complex64 (32) or complex128
 * (64): replaces STRUMPACK_factor (strumpack_finter.c:84).  Returns
 * WT_ERR_PIVOT when a diagonal tile had an exactly zero pivot. */
int wt_solver_factor(wt_solver *h, int factor_prec);

/* Pivot report of the last factorization in factor_prec.  Pivoting is
 * partial within 64-row diagonal tiles; magnitudes are |re| + |im| and each
 * ratio is taken against the tile's largest entry when its elimination
 * starts (after the updates of the earlier pivots).  A small pivot ratio or a
 * large growth flags a factor that refinement may not recover.  No
 * counterpart in the reference: STRUMPACK's shim passes only a failure code
 * to ierr (strumpack_finter.c:103-111).  WT_ERR_STATE before a
 * factorization. */
typedef struct wt_pivot_report {
    int32_t zero_pivots;     /* 1: an exactly zero pivot was replaced (the factorization gave WT_ERR_PIVOT) */
    int32_t tiny_pivots;     /* pivots with |d| <= 2^-12 (fp32) or 2^-26 (fp64) times max|tile| */
    double min_pivot_ratio;  /* min over pivots of |d| / max|tile| */
    double max_growth;       /* max over tiles of max|U| / max|tile| */
} wt_pivot_report;
int wt_solver_pivot_report(const wt_solver *h, int factor_prec, wt_pivot_report *out);
/* Tile-shape autotuner of the factorization: factors once per GEMM tile
 * variant with every launch timed, then keeps each launch's fastest variant
 * for the later factorizations of this handle (the handle is left
 * factored).  *speedup: GEMM time under the default choice over the tuned
 * time.  No counterpart in the reference. */
int wt_solver_autotune(wt_solver *h, int factor_prec, double *speedup);
/* The GEMM tile variant of every launch of the factorization (which = 0)
 * or of the block solve for nrhs columns (which = 1), so that a tuning can
 * be stored and applied again without re-measuring.  get: vars may be NULL
 * to query *count (plans are built on demand: the factorization's needs
 * assembled values); set: count must equal the plan's launch count.  No
 * counterpart in the reference. */
int wt_solver_get_tiles(wt_solver *h, int factor_prec, int which, int nrhs, int32_t *vars, int64_t *count);
/* Capture the factorization in a HIP graph on its first run and replay it
 * (default off for a single handle; profiling runs eagerly). */
int wt_solver_set_graphs(wt_solver *h, int enable);
/* complex64 GEMMs on the bf16 matrix cores (default on): the large launches
 * of the factorization and solve form each real fp32 product from a
 * three-way bf16 split of both operands (six partial products, fp32
 * accumulation), which keeps about the error of one fp32 rounding per
 * product; off keeps every product on the fp32 matrix cores.  Applies to
 * plans built later and maps the current ones' large tiles.  No
 * counterpart in the reference (STRUMPACK's complex64 GEMMs are fp32). */
int wt_solver_set_gemm_split(wt_solver *h, int enable);
/* Back substitution: the update products y_p -= U12 x_u of a tree level,
 * few output tiles with a long K near the root, are split along K into
 * pieces of at least min_depth (default 512) whose partial products are
 * summed in a fixed order; 0 keeps them whole.  The split depends on the
 * tree only, so a column is solved the same in any block.  Applies from the
 * next solve.  No counterpart in the reference (STRUMPACK's solve). */
int wt_solver_set_splitk(wt_solver *h, int min_depth);
int wt_solver_set_tiles(wt_solver *h, int factor_prec, int which, int nrhs, const int32_t *vars, int64_t count);
int wt_solver_release(wt_solver *h, int factor_prec);

/* Solve A X = B for nrhs right-hand sides (row-major n x nrhs complex128,
 * unknown-major: all columns of one unknown contiguous).  With the
 * complex64 factor each block of <= 64 columns is refined with FP64
 * residuals until max_j ||r_j||/||b_j|| <= tol or max_iter solves.
 * Refinement also stops when a step does not lower the residual.  A block
 * that ends above tol gives WT_ERR_NOCONV after every block has been solved
 * and X and *stats written (stats->converged = 0).  tol <= 0 asks for
 * exactly max_iter solves and never reports WT_ERR_NOCONV.
 * Replaces STRUMPACK_solve (strumpack_finter.c:102) for a whole block. */
int wt_solver_solve(wt_solver *h, int factor_prec, int nrhs, const void *b, void *x, int on_device,
                    double tol, int max_iter, wt_solve_stats *stats);

/* Y = A X with the solver's FP64 CSR values, row-major n x nrhs blocks
 * (host or device).  Used to form right-hand sides and residuals. */
int wt_solver_apply(wt_solver *h, int nrhs, const void *x, void *y, int on_device);

/* Device block d_x (row-major n x nrhs complex128) of homogeneous plane
 * waves u_j(r) = d_j exp(-i kmag d_j.r), d_j = dirs[3j..3j+2] (unit,
 * z positive down), r in metres on the solver's grid with spacing dx. */
int wt_solver_planewaves(wt_solver *h, int nrhs, const double *dirs, double kmag, double dx, void *d_x);

/* ---- plane-wave right-hand sides of a layered half-space ---- */

/* Thomson-Haskell depth profiles of plane waves incident from below on a
 * layered half-space; replaces HASKGRN (haskgrn.f:31-672), whose inputs are
 * module globals (HASKELL_MODULE vp1d, vs1d, rh1d, h1d, vp1disc, vs1disc,
 * nl; SRCPRM_MODULE ain, srctyp; MODEL_MODULE dz, nz, zorig), one source
 * per call there and ns sources per call here.
 *   nl layers (<= 64), the last one the half-space: vp, vs continuous speeds
 *   (they set the horizontal speed cv = v(nl)/sin(ain)), vpd, vsd the
 *   discrete speeds of the layer terms, rho densities, top = depth of each
 *   layer's top (top[0] <= zorig);
 *   ain[ns] incidence angles (rad), is_s[ns] = 1 for an incident S wave
 *   (srctyp 'ps'), 0 for P; omega angular frequency (the reference's df);
 *   depths dz (iz-1) + zorig, iz = 1..nz.
 * prec WT_PREC_F32: float arrays and scalars rounded to float, complex64
 * profiles, with the reference's single-precision subexpressions (as it is
 * compiled); WT_PREC_F64: double / complex128 throughout (the reference
 * with 8-byte REAL).
 * Out: device d_u, d_w [ns][nz] horizontal (along the propagation azimuth)
 * and vertical (positive down) profiles; device d_pic[ns] = omega / cv
 * (double); host status[ns]: 0, or 1 when a depth lies above top[0] (the
 * reference's ierr = 1; the profile is zero from that depth on).  nl < 2
 * with nz > 1 is refused (the reference would read h1d(0)).  Runs on the
 * given HIP stream (NULL: the default stream) and returns when it is done;
 * scratch is stream-ordered, so other streams are not waited for. */
int wt_haskgrn(int prec, int nl, const void *vp, const void *vs, const void *rho, const void *top,
               const void *vpd, const void *vsd, int ns, const void *ain, const int32_t *is_s, double omega,
               int nz, double dz, double zorig, void *d_u, void *d_w, double *d_pic, int32_t *status,
               void *stream);

/* wt_haskgrn with no device allocation.  No counterpart in the reference:
 * an entry point for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments, outputs and errors as wt_haskgrn, with at most
 * WT_HASKGRN_NOALLOC_MAX_SOURCES sources (more: WT_ERR_ARG and a message
 * on stderr).  The staged inputs and the status live in a fixed device
 * array of 67 KB per GPU, kept until the process ends; it belongs to the
 * current device, and the stream must be one of that device's.  Calls that
 * overlap on one GPU race on that array; calls made one after another are
 * safe (each returns when its work is done). */
#define WT_HASKGRN_NOALLOC_MAX_SOURCES 4096
int wt_haskgrn_noalloc(int prec, int nl, const void *vp, const void *vs, const void *rho, const void *top,
                       const void *vpd, const void *vsd, int ns, const void *ain, const int32_t *is_s,
                       double omega, int nz, double dz, double zorig, void *d_u, void *d_w, double *d_pic,
                       int32_t *status, void *stream);

/* 3-D background fields of plane waves from their depth profiles (the
 * expansion elwave's missing planewvexp.f90 would do; defined by this
 * library):
 *   u_j = (U_j(z) cos phi_j, U_j(z) sin phi_j, W_j(z))
 *         * exp(-i kh_j ((ix-1) dx cos phi_j + (iy-1) dx sin phi_j)),
 * phi_j the azimuth of horizontal propagation from +x towards +y (host,
 * radians), kh_j = d_kh[j] (device; wt_haskgrn's d_pic), profiles d_u, d_w
 * [nrhs][nz] (device, complex64 for prec WT_PREC_F32, complex128 for
 * WT_PREC_F64).  Out: device d_x, row-major n x nrhs complex128 in the
 * unknown order (node z fastest, then x, then y; components u, v, w).
 * Runs on the given stream (NULL: default) and returns when it is done. */
int wt_planewave_layered(int nx, int ny, int nz, double dx, int prec, int nrhs, const void *d_u,
                         const void *d_w, const double *d_kh, const double *phi, void *d_x, void *stream);

/* Scattered-field source; replaces HETFLDSRC (hetfldsrc.f90:17-1039):
 *   sfld_j = srcfn_j (A_bg - A) u_bg_j,
 * A the handle's FP64 operator from its last wt_solver_assemble, A_bg the
 * operator of the background model (lam_bg, mu_bg complex128, rho_bg
 * double; (nz,nx,ny) z fastest; host, or device with on_device) assembled
 * with the same parameters.  d_ubg, d_sfld: device, row-major n x nrhs
 * complex128; srcfn: host, nrhs complex128 (NULL: all 1).  Node numbering
 * is the unknown order of the CSR; the reference's own (hetfldsrc.f90:1020)
 * agrees with it only when ny = nz.  WT_ERR_STATE without an assembly. */
int wt_solver_scattered_source(wt_solver *h, const void *lam_bg, const void *mu_bg, const double *rho_bg,
                               int on_device, int nrhs, const void *d_ubg, const double *srcfn, void *d_sfld);

/* The same source without a solver handle: both operators (the model's
 * lam, mu, rho and the background's, complex128 / double, (nz,nx,ny) z
 * fastest; host, or device with on_device) assembled at *p into
 * stream-ordered scratch with the CSR pattern, then
 *   sfld_j = srcfn_j (A_bg - A) u_bg_j.
 * d_ubg, d_sfld: device, row-major with row strides ldu, ldo (<= 0: nrhs),
 * so the source can land in a column range of a wider block.  Scratch:
 * 2 nnz complex128 + the pattern.  Runs on the given stream and returns when
 * it is done. */
int wt_scattered_source_grid(const wt_grid_params *p, const void *lam, const void *mu, const double *rho,
                             const void *lam_bg, const void *mu_bg, const double *rho_bg, int on_device, int nrhs,
                             const void *d_ubg, int64_t ldu, const double *srcfn, void *d_sfld, int64_t ldo,
                             void *stream);

/* Total field from the scattered field: d_uh += d_ubg over count complex128
 * values (device); replaces the totfld branch of MAKEUH (makeuh.f:41-59).
 * Runs on the given stream (NULL: default) and returns when it is done. */
int wt_total_field(void *d_uh, const void *d_ubg, int64_t count, void *stream);

/* The HIP stream a solver handle runs its work on (for callers that put
 * their own device work in order with it). */
int wt_solver_stream(const wt_solver *h, void **stream);

/* ---- receivers ---- */

/* Receiver extraction compiled into a sparse matrix; replaces the receiver
 * and hydrophone loops of MAKEUH (makeuh.f:60-168) with SRCSETUP
 * (srcsetup.f:1-214, extraction branch) and SRCREG (srcreg.f), whose
 * inputs are module globals (RECPRM_MODULE xg, yg, zg, gspread, igreg,
 * receiver, recin; xr, yr, zr, rspread, irreg; MODEL_MODULE da, mu, omega,
 * dx, freesurf).
 *   ng receivers at (xg, yg, zg) [m from the first node], gspread, igreg
 *   (0: nearest node; 1-4: a (2 igreg + 1)^3 Kaiser-windowed sinc region
 *   for gspread <= 0.5, else an (INT(4 gspread) + 1)^3 discrete Dirac
 *   region, INT(4 gspread) + 1 <= 9), gfac: complex response per receiver
 *   (2 ng doubles) or NULL;
 *   nr hydrophones at (xr, yr, zr) with rspread, irreg: pressure
 *   (lambda + 2 mu / 3) div u at the nearest node, lam, mu complex128
 *   [nxnynz] (host), omega the angular frequency;
 *   freesurf[6] as MODEL_MODULE freesurf (top, +x, bottom, -x, +y, -y):
 *   taps beyond a free surface are mirrored with a change of sign, others
 *   dropped.
 * WT_ERR_ARG for a station whose taps would leave the grid. */
typedef struct wt_receivers wt_receivers;
int wt_receivers_create(int nx, int ny, int nz, double dx, const int32_t *freesurf, int ng, const double *xg,
                        const double *yg, const double *zg, double gspread, int igreg, const double *gfac,
                        int nr, const double *xr, const double *yr, const double *zr, double rspread,
                        int irreg, const void *lam, const void *mu, double omega_re, double omega_im,
                        wt_receivers **out);

/* Output rows: 3 ng (u, v, w per receiver: uest, vest, west) + nr (utest). */
int wt_receivers_rows(const wt_receivers *r);

/* d_out (rows x nrhs, row-major complex128, device) = E d_field, d_field a
 * solution block (n x nrhs row-major complex128, device; the total field
 * for MAKEUH's totfld).  Runs on the given stream and returns when done. */
int wt_receivers_extract(const wt_receivers *r, int nrhs, const void *d_field, void *d_out, void *stream);
int wt_receivers_destroy(wt_receivers *r);

/* Per-kernel timing with HIP events on the solver's stream: enable (1) or
 * disable (0) and reset.  Classes: 0 factor GEMMs on the 128 x 128 tile
 * (fp32 matrix cores), 1 solve GEMMs, 2 diagonal-tile LU, 3 extend-add,
 * 4 fills / value scatter, 5 assembly, 6 FP64 residual, 7 reorder /
 * gather, 8 super-panel triangular solves, 9 factor GEMMs on the other
 * tiles (64 x 64 bf16 splits included), 10 complex64 factor GEMMs as bf16
 * splits on the 128 x 128 tile (wt_solver_set_gemm_split).
 * _read returns launches, summed kernel milliseconds,
 * algorithmic flops (8 M N K per complex GEMM, 8 per residual complex
 * multiply-add) and algorithmic bytes.  No counterpart in the reference
 * (STRUMPACK's verbose statistics, strumpack_finter.c:32-33, are printed). */
int wt_solver_profile(wt_solver *h, int enable);
int wt_solver_profile_read(wt_solver *h, int cls, int64_t *launches, double *ms, double *flops,
                           double *bytes);

/* Drop-in for the reference's solver shim with the same argument list and
 * job semantics (strumpack_finter.c:9-120, interface.f90:297-307):
 * 1 analyse, 2 factor (complex64 values), 3 solve one right-hand side,
 * -1 release.  irptr/jcptr are 1-based and left unchanged. */
void strumpack_serial_finter(int *job, int *nrows, int *nnz, int *irptr, int *jcptr, void *a, void *rhs,
                             void *sol, int *ierr);
/* The reference's empty MPI stub (strumpack_finter.c:123-126). */
int strumpack_parallel_finter(void);

/* Batched complex GEMM of the front kernels on one problem, host arrays
 * (column-major): C = A B (mode 1), C -= A B (0), C += A B (2), C = -A B
 * (3; other modes are WT_ERR_ARG); optional
 * row permutation of B (bperm[k] = source row).  Runs `reps` times (C is
 * re-uploaded each time) and reports the last run's kernel time in *ms.
 * No counterpart in the reference: a test and measurement entry for the
 * kernel that replaces STRUMPACK's dense front updates. */
int wt_gemm_host(int prec, int M, int N, int K, const void *A, int64_t lda, const void *B, int64_t ldb,
                 void *C, int64_t ldc, int mode, const int32_t *bperm, int reps, double *ms);
/* The same with the tile variant chosen by the caller: 0 (128 x 128 fp32,
 * 128 x 64 fp64), 1 (256 x 64 fp32), 2 (64 x 64), 3 (128 x 64 fp32,
 * 64 x 128 fp64), or -1 for the default rule (fp32: 0 when N > 64 and the
 * problem has at least 1024 tiles of variant 0, else 2; fp64: 0).  No
 * counterpart in the reference. */
int wt_gemm_host_var(int prec, int var, int M, int N, int K, const void *A, int64_t lda, const void *B,
                     int64_t ldb, void *C, int64_t ldc, int mode, const int32_t *bperm, int reps, double *ms);

/* ---------------- batches of independent small models ---------------- */
/* The reference solves one model at one frequency per pass of its driver
 * loop: DISPERSION, ASMBLE, then STRUMPACK_SERIAL_FINTER jobs 2 and 3
 * (elwave.f90:147-186).  A batch runs that sequence for many models at once,
 * each with its own grid, medium and frequency: the models form one
 * block-diagonal system whose elimination forest is factored level by level
 * (every launch carries the fronts of all models at that depth), captured in
 * HIP graphs and replayed.  Unknowns of the batch are the models' unknowns
 * concatenated in model order (wt_batch_offsets). */
typedef struct wt_batch wt_batch;

/* dims: nmodels triples (nx, ny, nz).  Models are grouped, in order, so that
 * each group's CSR has at most max_group_nnz entries (<= 0: 2^31 - 1, the
 * 32-bit row-pointer limit) and, when max_group_bytes > 0, its numeric state
 * (complex64 factor, contribution blocks, 16-column solve and refinement
 * blocks, from the symbolic analysis) at most max_group_bytes; every group
 * is factored on its own stream. */
int wt_batch_create(int nmodels, const int32_t *dims, int leaf_nodes, int64_t max_group_nnz, int64_t max_group_bytes,
                    wt_batch **out);
int wt_batch_destroy(wt_batch *b);
/* info[8]: unknowns, CSR entries, fronts, levels (deepest tree), largest
 * front pivots, groups, factor entries, contribution-block entries;
 * *flops: the factorization's 8 (p^3/3 + p^2 u + p u^2) summed. */
int wt_batch_info(const wt_batch *b, int64_t *info, double *flops);
/* unk_off[nmodels + 1]: first unknown of each model; group_of[nmodels]. */
int wt_batch_offsets(const wt_batch *b, int64_t *unk_off, int32_t *group_of);
/* The HIP stream of the batch's first group (for callers that put their
 * own device work in order with the batch). */
int wt_batch_stream(const wt_batch *b, void **stream);
/* Capture the factorization and the refinement step in HIP graphs (default
 * on) or launch eagerly. */
int wt_batch_set_graphs(wt_batch *b, int enable);
/* MODEL_MODULE state of every model (as wt_solver_set_model), the models'
 * arrays concatenated in model order. */
int wt_batch_set_models(wt_batch *b, const void *lam, const void *mu, const double *rho, int on_device);
/* ASMBLE of every model with its own parameters params[nmodels] (grid sizes
 * must match the batch's; each carries its own frequency). */
int wt_batch_assemble(wt_batch *b, const wt_grid_params *params);
/* Numeric factorization of every model (job 2), WT_PREC_F32 or _F64. */
int wt_batch_factor(wt_batch *b, int factor_prec);
/* Tile-shape autotuner: factors once per GEMM tile variant, timing every
 * launch, and keeps each launch's fastest variant for later factorizations
 * (the batch is left factored).  *speedup: GEMM time under the default
 * choice over the tuned time.  No counterpart in the reference. */
int wt_batch_autotune(wt_batch *b, int factor_prec, double *speedup);
/* The union's symbolic analysis on the host only (no GPU): info[8] =
 * unknowns, CSR entries, fronts, levels, largest front pivots, factor
 * entries, contribution-block entries, solve-update entries. */
int wt_batch_symbolic_info(int nmodels, const int32_t *dims, int leaf_nodes, int64_t *info, double *flops);
/* The CSR pattern of SREL3D + SPARSE_ADJ2CRS (srgraphs.f:638-734,
 * sparse.f:13-49) computed on the host (wt_graph computes it on the GPU). */
int wt_graph_host(int nx, int ny, int nz, int base, int32_t *rowptr, int32_t *colidx);
/* Factor and solve group after group (arguments as wt_batch_solve),
 * releasing each group's factors before the next: only one group's numeric
 * state is resident, so a batch larger than device memory runs with groups
 * bounded by max_group_bytes. */
int wt_batch_run(wt_batch *b, int factor_prec, int nrhs, const int32_t *ncol, const void *rhs, void *x,
                 int on_device, double tol, int max_iter, wt_solve_stats *stats);
/* y = A x with every model's FP64 CSR values; x, y row-major (total
 * unknowns) x nrhs, nrhs <= 64. */
int wt_batch_apply(wt_batch *b, int nrhs, const void *x, void *y, int on_device);
/* Solve with FP64 iterative refinement (job 3 for many sources): rhs and x
 * are row-major (total unknowns) x nrhs, nrhs <= 64; model k uses its first
 * ncol[k] columns (the others are ignored on input and zero on output).
 * Refines until every model's columns reach |b - A x| / |b| <= tol or
 * max_iter solves, or no model's residual still falls; stats[nmodels] per
 * model (converged, stalled as for wt_solver_solve).  WT_ERR_NOCONV when
 * some model ends above tol > 0: x and stats are written for every model. */
int wt_batch_solve(wt_batch *b, int factor_prec, int nrhs, const int32_t *ncol, const void *rhs, void *x,
                   int on_device, double tol, int max_iter, wt_solve_stats *stats);

/* ---------------- device memory helpers ---------------- */
int wt_dev_alloc(size_t bytes, void **d_ptr);
int wt_dev_free(void *d_ptr);
int wt_h2d(void *d_dst, const void *src, size_t bytes);
int wt_d2h(void *dst, const void *d_src, size_t bytes);
int wt_dev_sync(void);
/* Device-to-device copy by a plain streaming kernel (bytes % 16 == 0): the
 * reference point for HBM traffic counters. */
int wt_dev_copy(void *d_dst, const void *d_src, size_t bytes);
/* rows x width bytes from d_src (row pitch spitch) to d_dst (row pitch
 * dpitch), device to device, on the given stream; returns when done. */
int wt_dev_copy2d(void *d_dst, size_t dpitch, const void *d_src, size_t spitch, size_t width, size_t rows,
                  void *stream);
/* Number of visible GPUs (0 without one). */
int wt_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* WTOMO3D_H */
