// Model pre-pass on the device: frequency-dependent complex moduli of an
// attenuating medium (dispersion.f:1-131, Aki & Richards eq. 5.88).
#include <hip/hip_runtime.h>

#include "wt_model.h"

namespace {

// Per node, with f = ln(omega / omega_base):
//   c = c_base max(1 + f q / pi, 0.1),  c~ = c (1 - i q / 2)  (q = 1/Q)
//   mu = vs~^2 rho,  lambda = vp~^2 rho - 2 mu
// (attenuate = 0: the base speeds unchanged, real moduli)
__global__ __launch_bounds__(256) void dispersion_kernel(int64_t nodes, const double* __restrict__ vp,
                                                         const double* __restrict__ vs,
                                                         const double* __restrict__ qpi,
                                                         const double* __restrict__ qsi,
                                                         const double* __restrict__ rho, double fac1, int attenuate,
                                                         cpx<double>* __restrict__ lam, cpx<double>* __restrict__ mu)
{
    const double pi = 3.141592653589793;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nodes; g += (int64_t)gridDim.x * blockDim.x) {
        const double qs = attenuate ? qsi[g] : 0.0, qp = attenuate ? qpi[g] : 0.0;
        double fs = 1.0 + (qs / pi) * fac1;
        if (fs < 0.1) fs = 0.1;
        const double sr = vs[g] * fs, si = -sr * qs / 2.0;
        const double mr = (sr * sr - si * si) * rho[g], mi = (sr * si + si * sr) * rho[g];
        double fp = 1.0 + (qp / pi) * fac1;
        if (fp < 0.1) fp = 0.1;
        const double pr = vp[g] * fp, pim = -pr * qp / 2.0;
        const double lr = (pr * pr - pim * pim) * rho[g] - 2.0 * mr;
        const double li = (pr * pim + pim * pr) * rho[g] - 2.0 * mi;
        mu[g] = cpx<double>(mr, mi);
        lam[g] = cpx<double>(lr, li);
    }
}

}  // namespace

int wt_launch_dispersion(int64_t nodes, const double* vp, const double* vs, const double* qpi, const double* qsi,
                         const double* rho, double fac1, int attenuate, cpx<double>* lam, cpx<double>* mu,
                         hipStream_t s)
{
    if (nodes <= 0) return WT_OK;
    const int64_t b = (nodes + 255) / 256;
    hipLaunchKernelGGL(dispersion_kernel, dim3((unsigned)(b > 65536 ? 65536 : b)), dim3(256), 0, s, nodes, vp, vs,
                       qpi, qsi, rho, fac1, attenuate, lam, mu);
    WT_HIP(hipGetLastError());
    return WT_OK;
}
