// C ABI: CSR pattern, assembly and device-memory helpers.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <vector>
#include "wt_common.h"
#include "wt_rt.h"

int64_t wt_nnz(int nx, int ny, int nz);
int wt_launch_pattern(int nx, int ny, int nz, int base, int32_t* d_rowptr, int32_t* d_colidx,
                      hipStream_t stream);
template <typename R>
int wt_launch_values(const wt_grid_params& p, const void* d_lam, const void* d_mu, const void* d_rho,
                     void* d_vals, int* d_err, hipStream_t stream);

// Resident device copy of the last matrix assembled through wt_asmble, read
// by the legacy solver shim (strumpack_serial_finter) so the values do not
// travel back over PCIe between the two calls.
namespace wt_resident {
void* d_vals = nullptr;
int64_t nnz = 0;
int nx = 0, ny = 0, nz = 0;
bool valid = false;
}  // namespace wt_resident

namespace {
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)rt_free(p); }
    int alloc(size_t n) { return rt_malloc(&p, n ? n : 1) == hipSuccess ? WT_OK : WT_ERR_NOMEM; }
};
}  // namespace

extern "C" {

int64_t wt_graph_nnz(int nx, int ny, int nz)
{
    if (nx < 1 || ny < 1 || nz < 1) return 0;
    return wt_nnz(nx, ny, nz);
}

int wt_graph_dev(int nx, int ny, int nz, int base, int32_t* d_rowptr, int32_t* d_colidx, void* stream)
{
    wt_rt_stream_guard g(stream);
    return wt_launch_pattern(nx, ny, nz, base, d_rowptr, d_colidx, (hipStream_t)stream);
}

int wt_graph(int nx, int ny, int nz, int base, int32_t* rowptr, int32_t* colidx)
{
    if (nx < 1 || ny < 1 || nz < 1 || !rowptr || !colidx) return WT_ERR_ARG;
    wt_rt_shared g;  // null-stream launch
    int64_t n = 3LL * nx * ny * nz, nnz = wt_nnz(nx, ny, nz);
    DevBuf r, c;
    if (r.alloc((n + 1) * 4) || c.alloc(nnz * 4)) return WT_ERR_NOMEM;
    int rc = wt_launch_pattern(nx, ny, nz, base, (int32_t*)r.p, (int32_t*)c.p, 0);
    if (rc) return rc;
    WT_HIP(rt_memcpy(rowptr, r.p, (n + 1) * 4, hipMemcpyDeviceToHost));
    WT_HIP(rt_memcpy(colidx, c.p, nnz * 4, hipMemcpyDeviceToHost));
    return WT_OK;
}

int wt_assemble_dev(const wt_grid_params* p, int prec, const void* d_lam, const void* d_mu,
                    const void* d_rho, void* d_vals, int* d_err, void* stream)
{
    if (!p) return WT_ERR_ARG;
    wt_rt_stream_guard g(stream);
    if (prec == WT_PREC_F32)
        return wt_launch_values<float>(*p, d_lam, d_mu, d_rho, d_vals, d_err, (hipStream_t)stream);
    if (prec == WT_PREC_F64)
        return wt_launch_values<double>(*p, d_lam, d_mu, d_rho, d_vals, d_err, (hipStream_t)stream);
    return WT_ERR_ARG;
}

int wt_assemble(const wt_grid_params* p, int prec, const void* lam, const void* mu, const void* rho,
                void* vals)
{
    if (!p || !lam || !mu || !rho || !vals) return WT_ERR_ARG;
    if (prec != WT_PREC_F32 && prec != WT_PREC_F64) return WT_ERR_ARG;
    wt_rt_shared g;  // null-stream launch
    int64_t nodes = (int64_t)p->nx * p->ny * p->nz, nnz = wt_nnz(p->nx, p->ny, p->nz);
    size_t rs = prec == WT_PREC_F32 ? 4 : 8;
    DevBuf l, m, r, v, e;
    if (l.alloc(nodes * 2 * rs) || m.alloc(nodes * 2 * rs) || r.alloc(nodes * rs) ||
        v.alloc(nnz * 2 * rs) || e.alloc(sizeof(int)))
        return WT_ERR_NOMEM;
    WT_HIP(rt_memcpy(l.p, lam, nodes * 2 * rs, hipMemcpyHostToDevice));
    WT_HIP(rt_memcpy(m.p, mu, nodes * 2 * rs, hipMemcpyHostToDevice));
    WT_HIP(rt_memcpy(r.p, rho, nodes * rs, hipMemcpyHostToDevice));
    WT_HIP(rt_memset(e.p, 0, sizeof(int)));
    int rc = wt_assemble_dev(p, prec, l.p, m.p, r.p, v.p, (int*)e.p, nullptr);
    if (rc) return rc;
    int herr = 0;
    WT_HIP(rt_memcpy(&herr, e.p, sizeof(int), hipMemcpyDeviceToHost));
    WT_HIP(rt_memcpy(vals, v.p, nnz * 2 * rs, hipMemcpyDeviceToHost));
    return herr ? WT_ERR_ARG : WT_OK;
}

void wt_asmble(int lunlog, int nzero, int nx, int ny, int nz, const int32_t* irptr, const int32_t* jcptr,
               void* a, int* ierr, const void* da, const void* mu, const float* rho, float omega_re,
               float omega_im, float dx, float pmld, float pmlf, int ipml, int fst, int usemin)
{
    (void)lunlog;
    (void)jcptr;
    wt_rt_shared g;  // null-stream launch
    *ierr = 0;
    int64_t n = 3LL * nx * ny * nz, nnz = wt_graph_nnz(nx, ny, nz);
    // Row-length check of asmble.f90:556-564 and total check :619-626.
    std::vector<int32_t> rp(n + 1);
    {
        std::vector<int32_t> ci(nnz);
        if (wt_graph(nx, ny, nz, 1, rp.data(), ci.data()) != WT_OK) { *ierr = 1; return; }
    }
    int bad = 0;
    for (int64_t i = 0; i < n; ++i)
        if (irptr[i + 1] - irptr[i] != rp[i + 1] - rp[i]) {
            if (bad < 10) printf("Error in asmble: nnz incorrect %lld\n", (long long)(i + 1));
            ++bad;
        }
    if ((int64_t)nzero != nnz) {
        printf("Predicted number of non-zeros is %d\nNumber of non-zeros found in matrix is %lld\n", nzero,
               (long long)nnz);
        *ierr = 1;
        return;
    }
    wt_grid_params p;
    p.nx = nx; p.ny = ny; p.nz = nz; p.ipml = ipml; p.fst = fst; p.usemin = usemin;
    p.dx = dx; p.pmld = pmld; p.pmlf = pmlf; p.omega_re = omega_re; p.omega_im = omega_im;
    int64_t nodes = (int64_t)nx * ny * nz;
    DevBuf l, m, r, e;
    if (l.alloc(nodes * 8) || m.alloc(nodes * 8) || r.alloc(nodes * 4) || e.alloc(sizeof(int))) {
        *ierr = 1;
        return;
    }
    if (wt_resident::d_vals && wt_resident::nnz != nnz) {
        (void)rt_free(wt_resident::d_vals);
        wt_resident::d_vals = nullptr;
    }
    if (!wt_resident::d_vals && rt_malloc(&wt_resident::d_vals, nnz * 8) != hipSuccess) {
        wt_resident::d_vals = nullptr;
        *ierr = 1;
        return;
    }
    wt_resident::valid = false;
    if (rt_memcpy(l.p, da, nodes * 8, hipMemcpyHostToDevice) != hipSuccess ||
        rt_memcpy(m.p, mu, nodes * 8, hipMemcpyHostToDevice) != hipSuccess ||
        rt_memcpy(r.p, rho, nodes * 4, hipMemcpyHostToDevice) != hipSuccess ||
        rt_memset(e.p, 0, sizeof(int)) != hipSuccess) {
        *ierr = 1;
        return;
    }
    if (wt_assemble_dev(&p, WT_PREC_F32, l.p, m.p, r.p, wt_resident::d_vals, (int*)e.p, nullptr) != WT_OK) {
        *ierr = 1;
        return;
    }
    int herr = 0;
    if (rt_memcpy(&herr, e.p, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        rt_memcpy(a, wt_resident::d_vals, nnz * 8, hipMemcpyDeviceToHost) != hipSuccess) {
        *ierr = 1;
        return;
    }
    if (herr) {
        printf("ERROR calling fdfd3d from ASBMLE\n");
        *ierr = 1;
        return;
    }
    if (bad) { *ierr = bad; return; }
    wt_resident::nnz = nnz;
    wt_resident::nx = nx; wt_resident::ny = ny; wt_resident::nz = nz;
    wt_resident::valid = true;
}

int wt_dev_alloc(size_t bytes, void** d_ptr)
{
    if (!d_ptr) return WT_ERR_ARG;
    return rt_malloc(d_ptr, bytes ? bytes : 1) == hipSuccess ? WT_OK : WT_ERR_NOMEM;
}
int wt_dev_free(void* d_ptr)
{
    WT_HIP(rt_free(d_ptr));
    return WT_OK;
}
int wt_h2d(void* d_dst, const void* src, size_t bytes)
{
    WT_HIP(rt_memcpy(d_dst, src, bytes, hipMemcpyHostToDevice));
    return WT_OK;
}
int wt_d2h(void* dst, const void* d_src, size_t bytes)
{
    WT_HIP(rt_memcpy(dst, d_src, bytes, hipMemcpyDeviceToHost));
    return WT_OK;
}
int wt_dev_copy2d(void* d_dst, size_t dpitch, const void* d_src, size_t spitch, size_t width, size_t rows,
                  void* stream)
{
    if (!d_dst || !d_src || width > dpitch || width > spitch) return WT_ERR_ARG;
    if (!width || !rows) return WT_OK;
    const hipStream_t st = (hipStream_t)stream;
    wt_rt_stream_guard g(stream);
    WT_HIP(hipMemcpy2DAsync(d_dst, dpitch, d_src, spitch, width, rows, hipMemcpyDeviceToDevice, st));
    WT_HIP(hipStreamSynchronize(st));
    return WT_OK;
}
int wt_dev_sync(void)
{
    WT_HIP(rt_device_sync());
    return WT_OK;
}
int wt_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"
