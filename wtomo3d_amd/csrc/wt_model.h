// Launcher of the model pre-pass kernels (wt_model.hip).
#pragma once
#include "wt_common.h"

// fac1 = ln(omega / omega_base); attenuate = 0 when omega or omega_base is
// not positive: the moduli of the base speeds (dispersion.f:72-75)
int wt_launch_dispersion(int64_t nodes, const double* vp, const double* vs, const double* qpi, const double* qsi,
                         const double* rho, double fac1, int attenuate, cpx<double>* lam, cpx<double>* mu,
                         hipStream_t s);
