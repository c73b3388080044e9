// C ABI of the plane-wave right-hand-side generator: Thomson-Haskell depth
// profiles (haskgrn.f:31-672) and their 3-D expansion.
#include <cmath>
#include <vector>

#include "wt_rhs.h"
#include "wt_rt.h"

int64_t wt_nnz(int nx, int ny, int nz);
int wt_launch_pattern(int nx, int ny, int nz, int base, int32_t* d_rowptr, int32_t* d_colidx, hipStream_t stream);
template <typename R>
int wt_launch_values(const wt_grid_params& p, const void* d_lam, const void* d_mu, const void* d_rho, void* d_vals,
                     int* d_err, hipStream_t stream);

namespace {

// Stream-ordered scratch: allocation and release are ordered on the caller's
// stream, so other streams' work is not waited for.
struct Buf {
    void* p = nullptr;
    hipStream_t s;
    explicit Buf(hipStream_t st) : s(st) {}
    ~Buf() { if (p) (void)rt_free_async(p, s); }
    int alloc(size_t n) { return rt_malloc_async(&p, n ? n : 1, s) == hipSuccess ? WT_OK : WT_ERR_NOMEM; }
};

template <typename RS>
int haskgrn_t(int nl, const void* const* layer, int ns, const void* ain, const int32_t* is_s, double omega, int nz,
              double dz, double zorig, void* d_u, void* d_w, double* d_pic, int32_t* status, hipStream_t st)
{
    std::vector<RS> lay((size_t)6 * nl);
    for (int k = 0; k < 6; ++k)
        for (int m = 0; m < nl; ++m) lay[(size_t)k * nl + m] = ((const RS*)layer[k])[m];
    Buf dl(st), da(st), ds(st), dst(st);
    if (dl.alloc(lay.size() * sizeof(RS)) || da.alloc(ns * sizeof(RS)) || ds.alloc(ns * 4) || dst.alloc(ns * 4))
        return WT_ERR_NOMEM;
    WT_HIP(hipMemcpyAsync(dl.p, lay.data(), lay.size() * sizeof(RS), hipMemcpyHostToDevice, st));
    WT_HIP(hipMemcpyAsync(da.p, ain, ns * sizeof(RS), hipMemcpyHostToDevice, st));
    WT_HIP(hipMemcpyAsync(ds.p, is_s, ns * 4, hipMemcpyHostToDevice, st));
    int rc = wt_launch_haskgrn<RS>(nl, (const RS*)dl.p, ns, (const RS*)da.p, (const int32_t*)ds.p, (RS)omega, nz,
                                   (RS)dz, (RS)zorig, (cpx<RS>*)d_u, (cpx<RS>*)d_w, d_pic, (int32_t*)dst.p, st);
    if (rc) {
        (void)hipStreamSynchronize(st);
        return rc;
    }
    WT_HIP(hipMemcpyAsync(status, dst.p, ns * 4, hipMemcpyDeviceToHost, st));
    WT_HIP(hipStreamSynchronize(st));
    return WT_OK;
}

}  // namespace

extern "C" {

int wt_haskgrn(int prec, int nl, const void* vp, const void* vs, const void* rho, const void* top, const void* vpd,
               const void* vsd, int ns, const void* ain, const int32_t* is_s, double omega, int nz, double dz,
               double zorig, void* d_u, void* d_w, double* d_pic, int32_t* status, void* stream)
{
    const hipStream_t st = (hipStream_t)stream;
    wt_rt_stream_guard g(stream);
    if (!vp || !vs || !rho || !top || !vpd || !vsd || !ain || !is_s || !d_u || !d_w || !d_pic || !status)
        return WT_ERR_ARG;
    if (nl < 1 || nl > WT_HASK_MAXL || ns < 1 || nz < 1) return WT_ERR_ARG;
    // a single layer has no layer above the half-space to walk through
    // (the reference would read h1d(0) for the second depth)
    if (nl < 2 && nz > 1) return WT_ERR_ARG;
    const void* layer[6] = {vp, vs, rho, top, vpd, vsd};
    if (prec == WT_PREC_F32)
        return haskgrn_t<float>(nl, layer, ns, ain, is_s, omega, nz, dz, zorig, d_u, d_w, d_pic, status, st);
    if (prec == WT_PREC_F64)
        return haskgrn_t<double>(nl, layer, ns, ain, is_s, omega, nz, dz, zorig, d_u, d_w, d_pic, status, st);
    return WT_ERR_ARG;
}

int wt_planewave_layered(int nx, int ny, int nz, double dx, int prec, int nrhs, const void* d_u, const void* d_w,
                         const double* d_kh, const double* phi, void* d_x, void* stream)
{
    const hipStream_t st = (hipStream_t)stream;
    wt_rt_stream_guard g(stream);
    if (nx < 1 || ny < 1 || nz < 1 || nrhs < 1 || !d_u || !d_w || !d_kh || !phi || !d_x) return WT_ERR_ARG;
    std::vector<double> cs((size_t)2 * nrhs);
    for (int j = 0; j < nrhs; ++j) {
        cs[2 * j] = std::cos(phi[j]);
        cs[2 * j + 1] = std::sin(phi[j]);
    }
    Buf dcs(st);
    if (dcs.alloc(cs.size() * 8)) return WT_ERR_NOMEM;
    WT_HIP(hipMemcpyAsync(dcs.p, cs.data(), cs.size() * 8, hipMemcpyHostToDevice, st));
    int rc;
    if (prec == WT_PREC_F32)
        rc = wt_launch_expand<float>((cpx<double>*)d_x, nx, ny, nz, dx, nrhs, (const cpx<float>*)d_u,
                                     (const cpx<float>*)d_w, d_kh, (const double*)dcs.p, st);
    else if (prec == WT_PREC_F64)
        rc = wt_launch_expand<double>((cpx<double>*)d_x, nx, ny, nz, dx, nrhs, (const cpx<double>*)d_u,
                                      (const cpx<double>*)d_w, d_kh, (const double*)dcs.p, st);
    else
        rc = WT_ERR_ARG;
    // the host copy of the angles must outlive the upload
    WT_HIP(hipStreamSynchronize(st));
    return rc;
}

int wt_scattered_source_grid(const wt_grid_params* p, const void* lam, const void* mu, const double* rho,
                             const void* lam_bg, const void* mu_bg, const double* rho_bg, int on_device, int nrhs,
                             const void* d_ubg, int64_t ldu, const double* srcfn, void* d_sfld, int64_t ldo,
                             void* stream)
{
    if (!p || !lam || !mu || !rho || !lam_bg || !mu_bg || !rho_bg || !d_ubg || !d_sfld || nrhs < 1) return WT_ERR_ARG;
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return WT_ERR_ARG;
    if (ldu <= 0) ldu = nrhs;
    if (ldo <= 0) ldo = nrhs;
    if (ldu < nrhs || ldo < nrhs) return WT_ERR_ARG;
    const hipStream_t st = (hipStream_t)stream;
    wt_rt_stream_guard g(stream);
    const int64_t nodes = (int64_t)p->nx * p->ny * p->nz, n = 3 * nodes, nnz = wt_nnz(p->nx, p->ny, p->nz);
    // stream-ordered scratch: the pattern, both operators, the models when
    // they come from the host, srcfn and the error word
    Buf drow(st), dcol(st), da(st), dab(st), dm(st), dsf(st), derr(st);
    if (drow.alloc((n + 1) * 4) || dcol.alloc(nnz * 4) || da.alloc(nnz * 16) || dab.alloc(nnz * 16) ||
        (!on_device && dm.alloc(nodes * 80)) || (srcfn && dsf.alloc((size_t)nrhs * 16)) || derr.alloc(4))
        return WT_ERR_NOMEM;
    const void *L = lam, *M = mu, *Lb = lam_bg, *Mb = mu_bg;
    const double *Rh = rho, *Rb = rho_bg;
    auto finish = [&](int rc) {
        const hipError_t e = hipStreamSynchronize(st);
        return rc ? rc : (e == hipSuccess ? WT_OK : WT_ERR_HIP);
    };
    if (!on_device) {
        char* m = (char*)dm.p;
        const void* src[6] = {lam, mu, lam_bg, mu_bg, rho, rho_bg};
        const size_t sz[6] = {16, 16, 16, 16, 8, 8};
        size_t off = 0;
        void* dst[6];
        for (int k = 0; k < 6; ++k) {
            dst[k] = m + off;
            if (hipMemcpyAsync(dst[k], src[k], nodes * sz[k], hipMemcpyHostToDevice, st) != hipSuccess)
                return finish(WT_ERR_HIP);
            off += nodes * sz[k];
        }
        L = dst[0]; M = dst[1]; Lb = dst[2]; Mb = dst[3];
        Rh = (const double*)dst[4]; Rb = (const double*)dst[5];
    }
    if ((srcfn && hipMemcpyAsync(dsf.p, srcfn, (size_t)nrhs * 16, hipMemcpyHostToDevice, st) != hipSuccess) ||
        hipMemsetAsync(derr.p, 0, 4, st) != hipSuccess)
        return finish(WT_ERR_HIP);
    int rc = wt_launch_pattern(p->nx, p->ny, p->nz, 0, (int32_t*)drow.p, (int32_t*)dcol.p, st);
    if (!rc) rc = wt_launch_values<double>(*p, L, M, Rh, da.p, (int*)derr.p, st);
    if (!rc) rc = wt_launch_values<double>(*p, Lb, Mb, Rb, dab.p, (int*)derr.p, st);
    const cpx<double>* u = (const cpx<double>*)d_ubg;
    const cpx<double>* sf = (const cpx<double>*)dsf.p;
    cpx<double>* out = (cpx<double>*)d_sfld;
    for (int c0 = 0; !rc && c0 < nrhs; c0 += 64) {
        const int nc = nrhs - c0 < 64 ? nrhs - c0 : 64;
        rc = wt_launch_scatsrc((const int32_t*)drow.p, (const int32_t*)dcol.p, (const cpx<double>*)da.p,
                               (const cpx<double>*)dab.p, u + c0, srcfn ? sf + c0 : nullptr, out + c0, n, nc, st,
                               ldu, ldo);
    }
    int he = 0;
    if (!rc && hipMemcpyAsync(&he, derr.p, 4, hipMemcpyDeviceToHost, st) != hipSuccess) rc = WT_ERR_HIP;
    rc = finish(rc);
    return rc == WT_OK && he ? WT_ERR_ARG : rc;
}

int wt_total_field(void* d_uh, const void* d_ubg, int64_t count, void* stream)
{
    if (!d_uh || !d_ubg || count < 0) return WT_ERR_ARG;
    const hipStream_t st = (hipStream_t)stream;
    wt_rt_stream_guard g(stream);
    int rc = wt_launch_add((cpx<double>*)d_uh, (const cpx<double>*)d_ubg, count, st);
    if (rc) return rc;
    WT_HIP(hipStreamSynchronize(st));
    return WT_OK;
}

}  // extern "C"
