// Launchers of the plane-wave right-hand-side kernels (wt_rhs.hip).
#pragma once
#include "wt_common.h"

#define WT_HASK_MAXL 64  // layers per model (the last one the half-space)

// d_layers: 6 nl-long arrays back to back: vp1d, vs1d, rh1d, h1d, vp1disc, vs1disc
template <typename RS>
int wt_launch_haskgrn(int nl, const RS* d_layers, int ns, const RS* d_ain, const int32_t* d_is_s, RS df, int nz,
                      RS dz, RS zorig, cpx<RS>* d_u, cpx<RS>* d_w, double* d_pic, int32_t* d_status,
                      hipStream_t s);

// d_csphi: (cos phi_j, sin phi_j) pairs; profiles source-major [nrhs][nz]
template <typename RP>
int wt_launch_expand(cpx<double>* d_x, int nx, int ny, int nz, double dx, int nrhs, const cpx<RP>* d_u,
                     const cpx<RP>* d_w, const double* d_kh, const double* d_csphi, hipStream_t s);

// u, out row-major with row strides ldu, ldo (<= 0: nrhs)
int wt_launch_scatsrc(const int32_t* rowptr, const int32_t* colidx, const cpx<double>* a, const cpx<double>* abg,
                      const cpx<double>* u, const cpx<double>* srcfn, cpx<double>* out, int64_t n, int nrhs,
                      hipStream_t s, int64_t ldu = 0, int64_t ldo = 0);

int wt_launch_add(cpx<double>* y, const cpx<double>* x, int64_t n, hipStream_t s);

int wt_launch_rows(const int32_t* rowptr, const int32_t* col, const cpx<double>* val, const cpx<double>* X,
                   cpx<double>* out, int rows, int nrhs, hipStream_t s);
