// Shared definitions for the wtomo3d MI355X library (HIP, gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include "wtomo3d.h"  // return codes WT_OK, WT_ERR_*

#define WT_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            fprintf(stderr, "wtomo3d: HIP error %s at %s:%d\n",                   \
                    hipGetErrorString(e_), __FILE__, __LINE__);                   \
            return WT_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

// Complex value with explicit component arithmetic, the operation order
// spelled out.  These operators are compiled with the default contraction:
// a product and the sum it feeds may fuse into one multiply-add (rounded
// once), so results can differ from a contraction-free evaluation by a few
// ulps.
template <typename R>
struct cpx {
    R re, im;
    __host__ __device__ cpx() {}
    __host__ __device__ constexpr cpx(R r, R i) : re(r), im(i) {}
};
template <typename R>
__host__ __device__ inline cpx<R> operator+(cpx<R> a, cpx<R> b) { return cpx<R>(a.re + b.re, a.im + b.im); }
template <typename R>
__host__ __device__ inline cpx<R> operator-(cpx<R> a, cpx<R> b) { return cpx<R>(a.re - b.re, a.im - b.im); }
template <typename R>
__host__ __device__ inline cpx<R> operator-(cpx<R> a) { return cpx<R>(-a.re, -a.im); }
template <typename R>
__host__ __device__ inline cpx<R> operator*(cpx<R> a, cpx<R> b)
{
    R re = a.re * b.re;
    re = re - a.im * b.im;
    R im = a.re * b.im;
    im = im + a.im * b.re;
    return cpx<R>(re, im);
}
template <typename R>
__host__ __device__ inline cpx<R> operator*(R s, cpx<R> a) { return cpx<R>(s * a.re, s * a.im); }
template <typename R>
__host__ __device__ inline cpx<R> operator/(cpx<R> a, R s) { return cpx<R>(a.re / s, a.im / s); }

// Smith's algorithm for complex division.
template <typename R>
__host__ __device__ inline cpx<R> cdiv(cpx<R> a, cpx<R> b)
{
    R br = b.re, bi = b.im;
    if ((br < 0 ? -br : br) >= (bi < 0 ? -bi : bi)) {
        R t = bi / br;
        R den = br + bi * t;
        return cpx<R>((a.re + a.im * t) / den, (a.im - a.re * t) / den);
    }
    R t = br / bi;
    R den = bi + br * t;
    return cpx<R>((a.re * t + a.im) / den, (a.im * t - a.re) / den);
}

