// wt_haskgrn_noalloc: the Thomson-Haskell profiles of wt_haskgrn without
// any device allocation.  wt_haskgrn stages the layer table, the incidence
// angles, the S flags and the per-source status through stream-ordered
// allocations on every call; here they live in one file-scope device array,
// `scratch`, of a fixed size.  Every GPU holds its own copy of the array
// until the process ends.  Its device address is looked up once per device
// with hipGetSymbolAddress (the first call on that device, with it current)
// and that address -- never the array's name, which in host code is only a
// host-side stand-in -- is what copies and kernel arguments use.
//
// All calls on one GPU share its array: calls that overlap on one GPU (from
// several host threads, or on several streams before one has returned) race
// on it.  A call returns only when its work is done, so calls made one
// after another from one thread are safe.  Each call writes every scratch
// entry it reads (the kernel stores the status of every source), so nothing
// has to start at zero.
#include <cstdio>
#include <mutex>

#include "wt_rhs.h"
#include "wt_rt.h"

namespace {

constexpr int kMaxSources = WT_HASKGRN_NOALLOC_MAX_SOURCES;
// byte offsets of the staged arrays (sized for double precision)
constexpr size_t kLayersOff = 0;  // 6 layer arrays of nl entries, back to back
constexpr size_t kAinOff = kLayersOff + 6 * WT_HASK_MAXL * sizeof(double);
constexpr size_t kIssOff = kAinOff + kMaxSources * sizeof(double);
constexpr size_t kStatusOff = kIssOff + kMaxSources * sizeof(int32_t);
constexpr size_t kScratchBytes = kStatusOff + kMaxSources * sizeof(int32_t);

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// The current device's copy of `scratch` (looked up on its first use).
int scratch_on_current_device(unsigned char** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return WT_ERR_HIP;
    if (dev < 0 || dev >= kMaxDevices) return WT_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "wtomo3d: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return WT_ERR_HIP;
        }
        g_scratch[dev] = (unsigned char*)p;
    }
    *out = g_scratch[dev];
    return WT_OK;
}

template <typename RS>
int profiles_noalloc(unsigned char* sc, int nl, const void* const* layer, int ns, const void* ain,
                     const int32_t* is_s, double omega, int nz, double dz, double zorig, void* d_u, void* d_w,
                     double* d_pic, int32_t* status, hipStream_t st)
{
    RS lay[6 * WT_HASK_MAXL];  // host staging on the stack
    for (int k = 0; k < 6; ++k)
        for (int m = 0; m < nl; ++m) lay[k * nl + m] = ((const RS*)layer[k])[m];
    RS* d_lay = (RS*)(sc + kLayersOff);
    RS* d_ain = (RS*)(sc + kAinOff);
    int32_t* d_iss = (int32_t*)(sc + kIssOff);
    int32_t* d_status = (int32_t*)(sc + kStatusOff);
    // every return waits for the queued work, which reads `lay` and the
    // caller's arrays
    auto finish = [&](int rc) {
        const hipError_t e = hipStreamSynchronize(st);
        return rc ? rc : (e == hipSuccess ? WT_OK : WT_ERR_HIP);
    };
    if (hipMemcpyAsync(d_lay, lay, (size_t)6 * nl * sizeof(RS), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_ain, ain, (size_t)ns * sizeof(RS), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_iss, is_s, (size_t)ns * 4, hipMemcpyHostToDevice, st) != hipSuccess)
        return finish(WT_ERR_HIP);
    int rc = wt_launch_haskgrn<RS>(nl, d_lay, ns, d_ain, d_iss, (RS)omega, nz, (RS)dz, (RS)zorig, (cpx<RS>*)d_u,
                                   (cpx<RS>*)d_w, d_pic, d_status, st);
    if (rc) return finish(rc);
    if (hipMemcpyAsync(status, d_status, (size_t)ns * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
        return finish(WT_ERR_HIP);
    return finish(WT_OK);
}

}  // namespace

extern "C" int wt_haskgrn_noalloc(int prec, int nl, const void* vp, const void* vs, const void* rho, const void* top,
                                  const void* vpd, const void* vsd, int ns, const void* ain, const int32_t* is_s,
                                  double omega, int nz, double dz, double zorig, void* d_u, void* d_w, double* d_pic,
                                  int32_t* status, void* stream)
{
    const hipStream_t st = (hipStream_t)stream;
    if (!vp || !vs || !rho || !top || !vpd || !vsd || !ain || !is_s || !d_u || !d_w || !d_pic || !status)
        return WT_ERR_ARG;
    if (nl < 1 || nl > WT_HASK_MAXL || ns < 1 || nz < 1) return WT_ERR_ARG;
    if (nl < 2 && nz > 1) return WT_ERR_ARG;
    if (prec != WT_PREC_F32 && prec != WT_PREC_F64) return WT_ERR_ARG;
    if (ns > kMaxSources) {
        fprintf(stderr, "wt_haskgrn_noalloc: %d sources exceed its scratch limit of %d (wt_haskgrn has none)\n", ns,
                kMaxSources);
        return WT_ERR_ARG;
    }
    wt_rt_stream_guard g(stream);
    unsigned char* sc = nullptr;
    int rc = scratch_on_current_device(&sc);
    if (rc) return rc;
    const void* layer[6] = {vp, vs, rho, top, vpd, vsd};
    if (prec == WT_PREC_F32)
        return profiles_noalloc<float>(sc, nl, layer, ns, ain, is_s, omega, nz, dz, zorig, d_u, d_w, d_pic, status,
                                       st);
    return profiles_noalloc<double>(sc, nl, layer, ns, ain, is_s, omega, nz, dz, zorig, d_u, d_w, d_pic, status, st);
}
