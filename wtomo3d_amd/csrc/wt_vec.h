// Vector kernels of the solve phase (layout changes, residual, norms).
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

#include "wt_common.h"

template <typename Rs, typename R>
int wt_to_elim(const cpx<Rs>* src, const int64_t* iperm, const double* scale, cpx<R>* y, int64_t n, int nrhs,
               hipStream_t s);
template <typename R, typename Rd>
int wt_from_elim(const cpx<R>* y, const int64_t* iperm, cpx<Rd>* dst, int64_t n, int nrhs, int accumulate,
                 hipStream_t s);
int64_t wt_residual_blocks(int64_t n);
template <typename Ra>
int wt_residual(const int32_t* rowptr, const int32_t* colidx, const cpx<Ra>* vals, const cpx<double>* x,
                const cpx<double>* b, cpx<double>* r, double* partial, double* rnorm2, int64_t n, int nrhs,
                hipStream_t s);
// The residual for 17..64 columns (n a multiple of 3) written straight as
// the next solve's input: y[perm[i] + c n] = (Ry) (scale[i] (b - A x)[i, c])
// (scale may be null); r itself is not stored.  rnorm2 gets |r_c|^2.
bool wt_residual_elim_ok(int64_t n, int nrhs);
template <typename Ra, typename Ry>
int wt_residual_elim(const int32_t* rowptr, const int32_t* colidx, const cpx<Ra>* vals, const cpx<double>* x,
                     const cpx<double>* b, const int64_t* perm, const double* scale, cpx<Ry>* y, double* partial,
                     double* rnorm2, int64_t n, int nrhs, hipStream_t s);
int wt_colnorm2(const cpx<double>* v, int64_t n, int nrhs, double* partial, int64_t nblk, double* out,
                hipStream_t s);
template <typename R>
int wt_copy(const cpx<R>* a, cpx<R>* b, int64_t n, hipStream_t s);
int wt_convert_c128_c64(const cpx<double>* a, cpx<float>* b, int64_t n, hipStream_t s);
int wt_widen_c64(const cpx<float>* a, cpx<double>* b, int64_t n, hipStream_t s);
int wt_planewave(cpx<double>* x, int nx, int ny, int nz, double dx, double kmag, const double* d_dirs, int nrhs,
                 hipStream_t s);
// Per-(segment, column) squared norms of a row-major n x nrhs block whose
// rows are cut at d_seg_off[0..nseg] (device): out[g * nrhs + c];
// partial holds nseg * nblk * nrhs doubles.
int wt_seg_colnorm2(const cpx<double>* v, const int64_t* d_seg_off, int nseg, int nrhs, double* partial, int nblk,
                    double* out, hipStream_t s);
// Zeroes columns c >= d_ncol[g] in the rows of every segment g.
int wt_seg_mask_cols(cpx<double>* v, const int64_t* d_seg_off, const int32_t* d_ncol, int nseg, int64_t n, int nrhs,
                     hipStream_t s);
