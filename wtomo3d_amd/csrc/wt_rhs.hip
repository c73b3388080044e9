// Plane-wave right-hand sides on the device: Thomson-Haskell depth profiles
// of a layered half-space (haskgrn.f:31-672), their expansion to 3-D
// background fields, and the scattered-field source (A_bg - A) u_bg of
// hetfldsrc.f90:17-1039.
#include <hip/hip_runtime.h>

#include "wt_rhs.h"

namespace {

// Per-layer terms kept in LDS: r1 r2 aag aag1 tt xh xh1 vsdisc e1 c01 s01 e2 c02 s02
constexpr int HK_T = 14;

// Propagator phase terms of one layer (haskgrn.f:168-187): trig when the
// horizontal speed exceeds the layer speed, hyperbolic otherwise.
__device__ __forceinline__ void hk_phase(double cv, double vlay, double z, double* q)
{
#pragma clang fp contract(off)
    if (cv > vlay) {
        q[0] = 1.0;
        q[1] = cos(z);
        q[2] = sin(z);
    } else {
        const double x = exp(z);
        q[0] = -1.0;
        q[1] = 0.5 * (x + 1.0 / x);
        q[2] = 0.5 * (x - 1.0 / x);
    }
}

// b <- A_m b over the two columns the surface vector (u, w, 0, 0) needs,
// with the alternating real/imaginary sign pattern of H53 p.23
// (haskgrn.f:231-321; fluid layer H53 eq 6.3).
__device__ __forceinline__ void hk_layer(const double* q, double* b)
{
#pragma clang fp contract(off)
    const double r1 = q[0], r2 = q[1], ag = q[2], ag1 = q[3], tm = q[4], x = q[5], y = q[6];
    const double cp = q[9], cq = q[12];
    const double pd = q[10] / r1, qd = q[13] / r2, pm = q[8] * q[10] * r1, qm = q[11] * q[13] * r2;
    double a11, a12, a13, a14, a21, a22, a23, a24, a31, a32, a33, a34, a41, a42, a43, a44;
    if (q[7] != 0.0) {
        a11 = ag * cp - ag1 * cq;
        a12 = ag1 * pd + ag * qm;
        a13 = (cq - cp) / tm;
        a14 = (pd + qm) / tm;
        a21 = -ag * pm - ag1 * qd;
        a22 = ag * cq - ag1 * cp;
        a23 = (pm + qd) / tm;
        a24 = a13;
        a31 = tm * ag * ag1 * (cp - cq);
        a32 = tm * (y * pd + x * qm);
        a33 = a22;
        a34 = a12;
        a41 = tm * (x * pm + y * qd);
        a42 = a31;
        a43 = a21;
        a44 = a11;
    } else {
        a11 = 0.0; a12 = -pd; a13 = -cp / tm; a14 = 0.0;
        a21 = 0.0; a22 = cp; a23 = pm / tm; a24 = 0.0;
        a31 = 0.0; a32 = tm * pd; a33 = cp; a34 = 0.0;
        a41 = 0.0; a42 = 0.0; a43 = 0.0; a44 = 0.0;
    }
    const double b11 = b[0], b12 = b[1], b21 = b[2], b22 = b[3], b31 = b[4], b32 = b[5], b41 = b[6], b42 = b[7];
    b[0] = a11 * b11 - a12 * b21 + a13 * b31 - a14 * b41;
    b[1] = a11 * b12 + a12 * b22 + a13 * b32 + a14 * b42;
    b[2] = a21 * b11 + a22 * b21 + a23 * b31 + a24 * b41;
    b[3] = -a21 * b12 + a22 * b22 - a23 * b32 + a24 * b42;
    b[4] = a31 * b11 - a32 * b21 + a33 * b31 - a34 * b41;
    b[5] = a31 * b12 + a32 * b22 + a33 * b32 + a34 * b42;
    b[6] = a41 * b11 + a42 * b21 + a43 * b31 + a44 * b41;
    b[7] = -a41 * b12 + a42 * b22 - a43 * b32 + a44 * b42;
}

// One block (one wave) per source.  Lanes build the per-layer terms; lane 0
// then forms the surface response and walks down the depth grid, updating
// the layer terms in the order the reference does: a depth at or below the
// half-space top is propagated with layer nl-1's terms (haskgrn.f:431-435),
// and on entering a new layer only one finished layer is restored to full
// thickness, its trig/hyperbolic choice taken from the speeds of the layer
// the previous depth's product loop ended past (haskgrn.f:447-476).
// RS is the reference's REAL: its single-precision subexpressions (the
// horizontal speed, layer thicknesses, half-space matrix constants, depths)
// are evaluated in RS, everything else in double.
template <typename RS>
__global__ __launch_bounds__(64) void haskgrn_kernel(int nl, const RS* __restrict__ vp1d, const RS* __restrict__ vs1d,
                                                     const RS* __restrict__ rh1d, const RS* __restrict__ h1d,
                                                     const RS* __restrict__ vp1disc, const RS* __restrict__ vs1disc,
                                                     const RS* __restrict__ ain, const int32_t* __restrict__ is_s,
                                                     RS df, int nz, RS dz, RS zorig, cpx<RS>* __restrict__ ugrn,
                                                     cpx<RS>* __restrict__ wgrn, double* __restrict__ pic_out,
                                                     int32_t* __restrict__ status)
{
#pragma clang fp contract(off)
    __shared__ double L[WT_HASK_MAXL * HK_T];
    __shared__ double h[WT_HASK_MAXL];
    const int src = blockIdx.x, lane = threadIdx.x;
    const bool S = is_s[src] != 0;
    const int N = nl - 1;
    // horizontal speed in REAL arithmetic (haskgrn.f:104-116); the sine of
    // the REAL angle is the double sine rounded to RS
    const RS sa = (RS)sin((double)ain[src]);
    const double cv = (double)(RS)((S ? vs1d[N] : vp1d[N]) / sa);
    const double pic = (double)df / cv;
    for (int m = lane; m < nl; m += 64) {
        double* q = L + HK_T * m;
        const double cova = cv / (double)vp1disc[m];
        const double covb = vs1disc[m] != (RS)0 ? cv / (double)vs1disc[m] : 1.0;
        double x = cova * cova - 1.0;
        q[0] = copysign(sqrt(fabs(x)), x);
        x = covb * covb - 1.0;
        q[1] = copysign(sqrt(fabs(x)), x);
        q[2] = 2.0 / (covb * covb);
        q[3] = q[2] - 1.0;
        q[7] = (double)vs1disc[m];
        if (m < N) {
            h[m] = (double)(RS)(h1d[m + 1] - h1d[m]);
            q[4] = (double)rh1d[m] * cv * cv;
            q[5] = q[2] * q[2];
            q[6] = q[3] * q[3];
            const double hc = pic * h[m];
            hk_phase(cv, (double)vp1disc[m], hc * q[0], q + 8);
            hk_phase(cv, (double)vs1disc[m], hc * q[1], q + 11);
        }
    }
    __syncthreads();
    if (lane != 0) return;
    cpx<RS>* U = ugrn + (int64_t)src * nz;
    cpx<RS>* W = wgrn + (int64_t)src * nz;
    pic_out[src] = pic;
    // scaled inverse E of the half-space (haskgrn.f:200-209)
    const double* qN = L + HK_T * N;
    const double xe = -qN[3] * cv * cv;
    const double e11 = (double)((RS)2 * vs1disc[N] * vs1disc[N]);
    const double e13 = (double)((RS)-1 / rh1d[N]);
    const double e22 = xe / qN[0], e24 = e13 / qN[0], e31 = xe / qN[1], e33 = -e13 / qN[1];
    const double e42 = -e11, e44 = e13;
    double b[8] = {1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0};
    for (int m = 0; m < N; ++m) hk_layer(L + HK_T * m, b);
    double c11 = e11 * b[0] + e13 * b[4], c12 = e11 * b[1] + e13 * b[5];
    double c21 = e22 * b[2] + e24 * b[6], c22 = e22 * b[3] + e24 * b[7];
    double c31 = e31 * b[0] + e33 * b[4], c32 = e31 * b[1] + e33 * b[5];
    double c41 = e42 * b[2] + e44 * b[6], c42 = e42 * b[3] + e44 * b[7];
    const double scl1 = (double)((RS)-1 / (vp1disc[N] * vp1disc[N]));
    const double scl2 = (double)((RS)-1 / ((RS)2 * vs1disc[N] * vs1disc[N]));
    c11 *= scl1; c12 *= scl1; c21 *= scl1; c22 *= scl1;
    c31 *= scl2; c32 *= scl2; c41 *= scl2; c42 *= scl2;
    const double x = -c11 * c42 + c21 * c32 + c31 * c22 - c41 * c12;
    const double y = c11 * c32 + c21 * c42 - c31 * c12 - c41 * c22;
    const double ag = x * x + y * y;
    double qrw, qiw, qru, qiu;
    if (!S) {
        const double tm = 2.0 * cv / ((double)vp1disc[N] * ag);
        qrw = tm * (c31 * x - c41 * y);
        qiw = tm * (-c31 * y - c41 * x);
        qru = tm * (c42 * x - c32 * y);
        qiu = tm * (-c42 * y - c32 * x);
    } else {
        const double tm = cv / ((double)vs1disc[N] * ag);
        qrw = tm * (-c11 * x + c21 * y);
        qiw = tm * (c11 * y + c21 * x);
        qru = tm * (-c22 * x + c12 * y);
        qiu = tm * (c22 * y + c12 * x);
    }
    U[0] = cpx<RS>((RS)qru, (RS)qiu);
    W[0] = cpx<RS>((RS)qrw, (RS)qiw);
    int ibovo = 1, ifin = 0, mstale = nl, rc = 0;
    for (int iz = 2; iz <= nz; ++iz) {
        const double dept = (double)(dz * (RS)(iz - 1) + zorig);
        int i = 2;
        while (i <= nl && !(dept < (double)h1d[i - 1])) ++i;
        if (i > nl) i = nl;
        const int ibov = i - 1;
        const double dloc = dept - (double)h1d[i - 2];
        if (dloc < 0.0) {
            rc = 1;
            for (int k = iz - 1; k < nz; ++k) {
                U[k] = cpx<RS>((RS)0, (RS)0);
                W[k] = cpx<RS>((RS)0, (RS)0);
            }
            break;
        }
        if (ibov != ibovo) {
            ibovo = ibov;
            if (ibov > 1) {
                ifin += 1;
                double* q = L + HK_T * (ifin - 1);
                const double hc = pic * h[ifin - 1];
                hk_phase(cv, (double)vp1disc[mstale - 1], hc * q[0], q + 8);
                hk_phase(cv, (double)vs1disc[mstale - 1], hc * q[1], q + 11);
            }
        }
        {
            double* q = L + HK_T * (ibov - 1);
            const double hc = pic * dloc;
            hk_phase(cv, (double)vp1disc[ibov - 1], hc * q[0], q + 8);
            hk_phase(cv, (double)vs1disc[ibov - 1], hc * q[1], q + 11);
        }
        double bb[8] = {1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0};
        for (int m = 0; m < ibov; ++m) hk_layer(L + HK_T * m, bb);
        mstale = ibov + 1;
        U[iz - 1] = cpx<RS>((RS)(bb[0] * qru - bb[1] * qiw), (RS)(bb[0] * qiu + bb[1] * qrw));
        W[iz - 1] = cpx<RS>((RS)(-bb[2] * qiu + bb[3] * qrw), (RS)(bb[2] * qru + bb[3] * qiw));
    }
    status[src] = rc;
}

// u_j(x, y, z) = (U_j(z) cos phi_j, U_j(z) sin phi_j, W_j(z)) e_j,
// e_j = exp(-i k_j ((ix-1) dx cos phi_j + (iy-1) dx sin phi_j)), written
// row-major n x nrhs in the unknown order (node z fastest, then x, then y).
template <typename RP>
__global__ __launch_bounds__(256) void expand_kernel(cpx<double>* __restrict__ x, int nx, int ny, int nz, double dx,
                                                     int nrhs, const cpx<RP>* __restrict__ U,
                                                     const cpx<RP>* __restrict__ Wp, const double* __restrict__ kh,
                                                     const double* __restrict__ cs_phi)
{
    const int64_t nodes = (int64_t)nx * ny * nz;
    const int64_t total = nodes * nrhs;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(t % nrhs);
        const int64_t g = t / nrhs;
        const int iz = (int)(g % nz), ix = (int)((g / nz) % nx), iy = (int)(g / ((int64_t)nz * nx));
        const double c = cs_phi[2 * j], s = cs_phi[2 * j + 1];
        const double ph = -kh[j] * dx * ((double)ix * c + (double)iy * s);
        double sn, cn;
        sincos(ph, &sn, &cn);
        const cpx<RP> u = U[(int64_t)j * nz + iz], w = Wp[(int64_t)j * nz + iz];
        const double ur = (double)u.re * cn - (double)u.im * sn, ui = (double)u.re * sn + (double)u.im * cn;
        const double wr = (double)w.re * cn - (double)w.im * sn, wi = (double)w.re * sn + (double)w.im * cn;
        cpx<double>* row = x + 3 * g * nrhs + j;
        row[0] = cpx<double>(ur * c, ui * c);
        row[nrhs] = cpx<double>(ur * s, ui * s);
        row[2 * nrhs] = cpx<double>(wr, wi);
    }
}

// sfld = srcfn_c sum_e (Abg_e - A_e) u_bg[col_e]: one wave per matrix row,
// lane = right-hand side, rows spread over the XCDs in contiguous ranges.
// u and out are row-major with row strides ldu and ldo (>= nrhs).
constexpr int SS_ROWS = 4;

__global__ __launch_bounds__(256) void scatsrc_kernel(const int32_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ colidx,
                                                      const cpx<double>* __restrict__ a,
                                                      const cpx<double>* __restrict__ abg,
                                                      const cpx<double>* __restrict__ u,
                                                      const cpx<double>* __restrict__ srcfn,
                                                      cpx<double>* __restrict__ out, int64_t n, int nrhs,
                                                      int64_t ldu, int64_t ldo)
{
    const int wave = threadIdx.x >> 6, c = threadIdx.x & 63;
    int64_t blk = blockIdx.x;
    {
        const int64_t nb = gridDim.x, q = nb / 8, r = nb % 8, xx = blk % 8, i = blk / 8;
        blk = xx * q + (xx < r ? xx : r) + i;
    }
    const int64_t row = blk * SS_ROWS + wave;
    if (row >= n || c >= nrhs) return;
    double ar = 0.0, ai = 0.0;
    const int32_t e1 = rowptr[row + 1];
    for (int32_t e = rowptr[row]; e < e1; ++e) {
        const cpx<double> d(abg[e].re - a[e].re, abg[e].im - a[e].im);
        const cpx<double> xv = u[(int64_t)colidx[e] * ldu + c];
        ar += d.re * xv.re - d.im * xv.im;
        ai += d.re * xv.im + d.im * xv.re;
    }
    const cpx<double> f = srcfn ? srcfn[c] : cpx<double>(1.0, 0.0);
    out[row * ldo + c] = cpx<double>(f.re * ar - f.im * ai, f.re * ai + f.im * ar);
}

// out[r, j] = sum_e val_e X[col_e, j] for a short list of sparse rows (the
// receiver extraction): one wave per row, lanes over the right-hand sides.
__global__ __launch_bounds__(256) void rows_kernel(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                   const cpx<double>* __restrict__ val,
                                                   const cpx<double>* __restrict__ X, cpx<double>* __restrict__ out,
                                                   int rows, int nrhs)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= rows) return;
    const int32_t e0 = rowptr[row], e1 = rowptr[row + 1];
    for (int j = lane; j < nrhs; j += 64) {
        double ar = 0.0, ai = 0.0;
        for (int32_t e = e0; e < e1; ++e) {
            const cpx<double> a = val[e], x = X[(int64_t)col[e] * nrhs + j];
            ar += a.re * x.re - a.im * x.im;
            ai += a.re * x.im + a.im * x.re;
        }
        out[(int64_t)row * nrhs + j] = cpx<double>(ar, ai);
    }
}

// u_h += u_bg (makeuh.f:41-59 with totfld)
__global__ __launch_bounds__(256) void add_kernel(cpx<double>* __restrict__ y, const cpx<double>* __restrict__ x,
                                                  int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = cpx<double>(y[i].re + x[i].re, y[i].im + x[i].im);
}

unsigned grid_for(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

}  // namespace

template <typename RS>
int wt_launch_haskgrn(int nl, const RS* d_layers, int ns, const RS* d_ain, const int32_t* d_is_s, RS df, int nz,
                      RS dz, RS zorig, cpx<RS>* d_u, cpx<RS>* d_w, double* d_pic, int32_t* d_status, hipStream_t s)
{
    if (nl < 1 || nl > WT_HASK_MAXL || ns < 1 || nz < 1) return WT_ERR_ARG;
    const RS* l = d_layers;
    hipLaunchKernelGGL(haskgrn_kernel<RS>, dim3(ns), dim3(64), 0, s, nl, l, l + nl, l + 2 * nl, l + 3 * nl,
                       l + 4 * nl, l + 5 * nl, d_ain, d_is_s, df, nz, dz, zorig, d_u, d_w, d_pic, d_status);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template <typename RP>
int wt_launch_expand(cpx<double>* d_x, int nx, int ny, int nz, double dx, int nrhs, const cpx<RP>* d_u,
                     const cpx<RP>* d_w, const double* d_kh, const double* d_csphi, hipStream_t s)
{
    const int64_t total = (int64_t)nx * ny * nz * nrhs;
    hipLaunchKernelGGL(expand_kernel<RP>, dim3(grid_for(total)), dim3(256), 0, s, d_x, nx, ny, nz, dx, nrhs, d_u, d_w,
                       d_kh, d_csphi);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

int wt_launch_scatsrc(const int32_t* rowptr, const int32_t* colidx, const cpx<double>* a, const cpx<double>* abg,
                      const cpx<double>* u, const cpx<double>* srcfn, cpx<double>* out, int64_t n, int nrhs,
                      hipStream_t s, int64_t ldu, int64_t ldo)
{
    if (ldu <= 0) ldu = nrhs;
    if (ldo <= 0) ldo = nrhs;
    if (nrhs < 1 || nrhs > 64 || ldu < nrhs || ldo < nrhs) return WT_ERR_ARG;
    const int64_t blocks = (n + SS_ROWS - 1) / SS_ROWS;
    hipLaunchKernelGGL(scatsrc_kernel, dim3((unsigned)blocks), dim3(256), 0, s, rowptr, colidx, a, abg, u, srcfn, out,
                       n, nrhs, ldu, ldo);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

int wt_launch_rows(const int32_t* rowptr, const int32_t* col, const cpx<double>* val, const cpx<double>* X,
                   cpx<double>* out, int rows, int nrhs, hipStream_t s)
{
    if (rows <= 0) return WT_OK;
    hipLaunchKernelGGL(rows_kernel, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s, rowptr, col, val, X, out, rows,
                       nrhs);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

int wt_launch_add(cpx<double>* y, const cpx<double>* x, int64_t n, hipStream_t s)
{
    if (n <= 0) return WT_OK;
    hipLaunchKernelGGL(add_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, x, n);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template int wt_launch_haskgrn<float>(int, const float*, int, const float*, const int32_t*, float, int, float, float,
                                      cpx<float>*, cpx<float>*, double*, int32_t*, hipStream_t);
template int wt_launch_haskgrn<double>(int, const double*, int, const double*, const int32_t*, double, int, double,
                                       double, cpx<double>*, cpx<double>*, double*, int32_t*, hipStream_t);
template int wt_launch_expand<float>(cpx<double>*, int, int, int, double, int, const cpx<float>*, const cpx<float>*,
                                     const double*, const double*, hipStream_t);
template int wt_launch_expand<double>(cpx<double>*, int, int, int, double, int, const cpx<double>*,
                                      const cpx<double>*, const double*, const double*, hipStream_t);
