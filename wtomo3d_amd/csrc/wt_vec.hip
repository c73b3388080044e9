// Vector kernels of the solve phase: reordering between the caller's
// layout and the elimination order, the FP64 CSR residual for iterative
// refinement, and per-right-hand-side norms.
//
// Caller blocks are row-major n x nrhs (all right-hand sides of one unknown
// contiguous), which makes the CSR residual read each matrix entry once for
// the whole block.  Inside the factor/solve the block is column-major in
// elimination order so the dense front solves are plain GEMMs.
#include "wt_vec.h"

#include <type_traits>

namespace {

constexpr int TP = 32;  // transpose tile

// y[i' + c n] = (R) (scale[o] src[o * nrhs + c]),  o = iperm[i']
template <typename Rs, typename R>
__global__ __launch_bounds__(256) void to_elim_kernel(const cpx<Rs>* __restrict__ src, const int64_t* __restrict__ iperm,
                                                      const double* __restrict__ scale, cpx<R>* __restrict__ y,
                                                      int64_t n, int nrhs)
{
    __shared__ cpx<R> t[TP][TP + 1];
    const int64_t i0 = (int64_t)blockIdx.x * TP;
    const int c0 = blockIdx.y * TP;
    const int tx = threadIdx.x % TP, ty = threadIdx.x / TP;  // 32 x 8
    for (int k = ty; k < TP; k += 8) {
        int64_t i = i0 + k;
        int c = c0 + tx;
        if (i < n && c < nrhs) {
            const int64_t o = iperm[i];
            const double sc = scale ? scale[o] : 1.0;
            cpx<Rs> v = src[o * nrhs + c];
            t[k][tx] = cpx<R>((R)(sc * (double)v.re), (R)(sc * (double)v.im));
        }
    }
    __syncthreads();
    for (int k = ty; k < TP; k += 8) {
        int c = c0 + k;
        int64_t i = i0 + tx;
        if (i < n && c < nrhs) y[(int64_t)c * n + i] = t[tx][k];
    }
}

// dst[iperm[i'] * nrhs + c] (+)= y[i' + c n]
template <typename R, typename Rd>
__global__ __launch_bounds__(256) void from_elim_kernel(const cpx<R>* __restrict__ y, const int64_t* __restrict__ iperm,
                                                        cpx<Rd>* __restrict__ dst, int64_t n, int nrhs, int accumulate)
{
    __shared__ cpx<R> t[TP][TP + 1];
    const int64_t i0 = (int64_t)blockIdx.x * TP;
    const int c0 = blockIdx.y * TP;
    const int tx = threadIdx.x % TP, ty = threadIdx.x / TP;
    for (int k = ty; k < TP; k += 8) {
        int c = c0 + k;
        int64_t i = i0 + tx;
        if (i < n && c < nrhs) t[tx][k] = y[(int64_t)c * n + i];
    }
    __syncthreads();
    for (int k = ty; k < TP; k += 8) {
        int64_t i = i0 + k;
        int c = c0 + tx;
        if (i < n && c < nrhs) {
            cpx<Rd>* p = dst + iperm[i] * nrhs + c;
            cpx<R> v = t[k][tx];
            if (accumulate) *p = cpx<Rd>(p->re + (Rd)v.re, p->im + (Rd)v.im);
            else *p = cpx<Rd>((Rd)v.re, (Rd)v.im);
        }
    }
}

// r = b - A x for a row-major block of up to 64 columns: one wave per
// matrix row, lane = right-hand side, so each matrix entry is read once for
// the whole block and the x rows are read as contiguous 64-wide segments.
// Per-block partial sums of |r|^2 per column go to `partial`.
constexpr int SP_ROWS = 4;  // rows per block (one per wave)
constexpr int RES_TB = 3;  // triplets per batch of the node-wise residual

// acc -= a x (SUB) or acc += a x as four fused multiply-adds in a fixed
// order, so that every residual kernel forms each term identically.  The
// signs are operand modifiers: a complex term costs four FP64 operations,
// the least there is -- with 64 right-hand sides these kernels do as much
// FP64 arithmetic per byte of matrix as the vector units can keep up with.
template <bool SUB, typename Ra>
__device__ __forceinline__ void res_term(cpx<double>& acc, const cpx<Ra> a, const cpx<double> xv)
{
    const double ar = SUB ? -(double)a.re : (double)a.re, ai = SUB ? -(double)a.im : (double)a.im;
    acc.re = fma(ar, xv.re, acc.re);
    acc.re = fma(-ai, xv.im, acc.re);
    acc.im = fma(ar, xv.im, acc.im);
    acc.im = fma(ai, xv.re, acc.im);
}

// r = b - A x (or A x when b is null) for a row-major block of nrhs <= CW
// columns: each wave handles 64 / CW rows, CW lanes per row (one per
// column), so few right-hand sides still fill the wave.  Every row of the
// elastic CSR holds its entries in triplets (the three components of one
// neighbour node, consecutive columns), so one column index serves three
// entries; the terms are added in the same order as one entry at a time.
template <typename Ra, int CW, bool SUB>
__global__ __launch_bounds__(256) void residual_kernel(const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ colidx,
                                                       const cpx<Ra>* __restrict__ vals,
                                                       const cpx<double>* __restrict__ x,
                                                       const cpx<double>* __restrict__ b, cpx<double>* __restrict__ r,
                                                       double* __restrict__ partial, int64_t n, int nrhs)
{
    constexpr int RPW = 64 / CW;
    __shared__ double s_part[SP_ROWS * 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, sub = lane / CW, c = lane % CW;
    // each of the 8 XCDs gets a contiguous range of rows, so the x rows of
    // 27-point neighbours are shared through its L2
    int64_t blk = blockIdx.x;
    {
        const int64_t nb = gridDim.x, q = nb / 8, rr = nb % 8, xx = blk % 8, i = blk / 8;
        blk = xx * q + (xx < rr ? xx : rr) + i;
    }
    const int64_t row = (blk * SP_ROWS + wave) * RPW + sub;
    double nrm = 0.0;
    if (row < n && c < nrhs) {
        cpx<double> acc = SUB ? b[row * nrhs + c] : cpx<double>(0.0, 0.0);
        auto term = [&](const cpx<Ra> a, const cpx<double> xv) { res_term<SUB>(acc, a, xv); };
        const int32_t e0 = rowptr[row], e1 = rowptr[row + 1];
        if ((e1 - e0) % 3 == 0) {
            for (int32_t e = e0; e < e1; e += 3) {
                const cpx<double>* xp = x + (int64_t)colidx[e] * nrhs + c;
                const cpx<Ra> a0 = vals[e], a1 = vals[e + 1], a2 = vals[e + 2];
                const cpx<double> x0 = xp[0], x1 = xp[nrhs], x2 = xp[2 * nrhs];
                term(a0, x0);
                term(a1, x1);
                term(a2, x2);
            }
        } else {
            for (int32_t e = e0; e < e1; ++e) term(vals[e], x[(int64_t)colidx[e] * nrhs + c]);
        }
        r[row * nrhs + c] = acc;
        nrm = acc.re * acc.re + acc.im * acc.im;
    }
    s_part[threadIdx.x] = nrm;
    __syncthreads();
    if ((int)threadIdx.x < nrhs) {
        double t = 0.0;
        for (int w = 0; w < SP_ROWS; ++w)
            for (int k = 0; k < RPW; ++k) t += s_part[w * 64 + k * CW + threadIdx.x];
        partial[blk * nrhs + threadIdx.x] = t;
    }
}

// The same residual for blocks of more than 16 columns, one wave per node
// (its three rows), lane = column.  The three rows of a node have the same
// columns, so each neighbour node's x rows (three 64-wide segments) are
// loaded once for the node instead of once per row.  The wave loads its
// node's 3 x len matrix values (up to 243) and len / 3 column indices into
// LDS with a few coalesced vector loads, then walks the triplets TB at a
// time, reading the values as LDS broadcasts.  Each row adds its terms in
// the row kernel's order.
//
// Blocks are mapped to XCDs in contiguous node ranges (the hardware deals
// workgroups to the 8 XCDs round-robin), so the x rows a node's 27
// neighbours share are fetched into one XCD's L2 instead of all eight.
//
// With `y` set, the kernel writes the solve's next input instead of r:
// y[perm[i] + c n] = (Ry) (scale[i] r[i, c]) -- the row-equilibrated
// residual in elimination order, column-major -- and r is not stored.  The
// three rows of a node are consecutive in elimination order, so each lane
// stores 3 consecutive entries of its column.
//
// Measured at 96x96x64, 64 columns (A x): 3.30 ms; the earlier form with
// the values through scalar loads two triplets at a time took 3.43 ms, one
// with them loaded per lane 7.0 ms.  Per node the wave reads 81 x rows of
// 1 KB through L1 / L2 (49 GB per call, 15 TB/s at 3.3 ms) and does 972
// FP64 fused multiply-adds per lane (0.93 ms at the FP64 peak).
template <typename Ra, typename Ry, bool SUB, int TB>
__global__ __launch_bounds__(256) void residual_node_kernel(const int32_t* __restrict__ rowptr,
                                                                const int32_t* __restrict__ colidx,
                                                                const cpx<Ra>* __restrict__ vals,
                                                                const cpx<double>* __restrict__ x,
                                                                const cpx<double>* __restrict__ b,
                                                                cpx<double>* __restrict__ r,
                                                                const int64_t* __restrict__ perm,
                                                                const double* __restrict__ scale,
                                                                cpx<Ry>* __restrict__ y,
                                                                double* __restrict__ partial, int64_t nnode, int nrhs)
{
    constexpr int MAXE = 3 * 81;
    __shared__ cpx<Ra> s_val[4][MAXE];
    __shared__ int32_t s_col[4][27];
    __shared__ double s_part[4 * 64];
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int c = threadIdx.x & 63;
    int64_t blk = blockIdx.x;
    {
        const int64_t nb = gridDim.x, q = nb / 8, rr = nb % 8, xx = blk % 8, i = blk / 8;
        blk = xx * q + (xx < rr ? xx : rr) + i;
    }
    const int64_t node = blk * 4 + wave;
    double nrm = 0.0;
    if (node < nnode) {  // wave-uniform
        const bool lane = c < nrhs;
        const int cc = lane ? c : 0;  // idle lanes read column 0 and store nothing
        const int64_t row0 = 3 * node;
        cpx<double> acc[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] = (SUB && lane) ? b[(row0 + k) * nrhs + cc] : cpx<double>(0.0, 0.0);
        auto term = [&](cpx<double>& a_, const cpx<Ra> a, const cpx<double> xv) { res_term<SUB>(a_, a, xv); };
        const int32_t e0 = rowptr[row0], e1 = rowptr[row0 + 1], e2 = rowptr[row0 + 2], e3 = rowptr[row0 + 3];
        const int32_t len = e1 - e0;
        if (len == e2 - e1 && len == e3 - e2 && len % 3 == 0 && len <= 81) {
            const int nt = len / 3;
            for (int i = c; i < 3 * len; i += 64) s_val[wave][i] = vals[e0 + i];
            if (c < nt) s_col[wave][c] = colidx[e0 + 3 * c];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const cpx<Ra>* sv = s_val[wave];
            int t0 = 0;
            auto batch = [&](auto tb_c) {
                constexpr int B = decltype(tb_c)::value;
                cpx<double> xv[B][3];
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    const cpx<double>* xp = x + (int64_t)s_col[wave][t0 + u] * nrhs + cc;
                    xv[u][0] = xp[0];
                    xv[u][1] = xp[nrhs];
                    xv[u][2] = xp[2 * nrhs];
                }
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    const int e = 3 * (t0 + u);
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        term(acc[k], sv[k * len + e], xv[u][0]);
                        term(acc[k], sv[k * len + e + 1], xv[u][1]);
                        term(acc[k], sv[k * len + e + 2], xv[u][2]);
                    }
                }
                t0 += B;
            };
            while (t0 + TB <= nt) batch(std::integral_constant<int, TB>());
            while (t0 < nt) batch(std::integral_constant<int, 1>());
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k)
                for (int32_t e = rowptr[row0 + k]; e < rowptr[row0 + k + 1]; ++e)
                    term(acc[k], vals[e], x[(int64_t)colidx[e] * nrhs + cc]);
        }
        if (lane) {
#pragma unroll
            for (int k = 0; k < 3; ++k) nrm += acc[k].re * acc[k].re + acc[k].im * acc[k].im;
            if (y) {
                const int64_t n = 3 * nnode;
                cpx<Ry>* yc = y + (int64_t)c * n;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double sc = scale ? scale[row0 + k] : 1.0;
                    yc[perm[row0 + k]] = cpx<Ry>((Ry)(sc * acc[k].re), (Ry)(sc * acc[k].im));
                }
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) r[(row0 + k) * nrhs + c] = acc[k];
            }
        }
    }
    s_part[threadIdx.x] = nrm;
    __syncthreads();
    if ((int)threadIdx.x < nrhs) {
        double t = 0.0;
        for (int w = 0; w < 4; ++w) t += s_part[w * 64 + threadIdx.x];
        partial[blk * nrhs + threadIdx.x] = t;
    }
}

// sum over blocks of partial[blk * nrhs + c] -> out[c] (fixed order)
__global__ void reduce_cols_kernel(const double* __restrict__ partial, int64_t nblk, int nrhs,
                                   double* __restrict__ out)
{
    __shared__ double s[256];
    const int c = blockIdx.x;
    double acc = 0.0;
    for (int64_t b = threadIdx.x; b < nblk; b += 256) acc += partial[b * nrhs + c];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = s[0];
}

// per-column |v|^2 of a row-major block, partial sums per block: block b
// sums its share of the rows, thread t column t % CW of every (256 / CW)-th
// row, so each pass reads whole rows (contiguous) of the block
__global__ __launch_bounds__(256) void colnorm_kernel(const cpx<double>* __restrict__ v, int64_t n, int nrhs, int cw,
                                                      double* __restrict__ partial)
{
    __shared__ double s[256];
    const int c = threadIdx.x % cw, r0t = threadIdx.x / cw, rstep = 256 / cw;
    double acc = 0.0;
    const int64_t rows_per_blk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_blk;
    const int64_t r1 = r0 + rows_per_blk < n ? r0 + rows_per_blk : n;
    if (c < nrhs)
        for (int64_t i = r0 + r0t; i < r1; i += rstep) {
            const cpx<double> a = v[i * nrhs + c];
            acc += a.re * a.re + a.im * a.im;
        }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off >= cw; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if ((int)threadIdx.x < nrhs) partial[(int64_t)blockIdx.x * nrhs + threadIdx.x] = s[threadIdx.x];
}

template <typename R>
__global__ void copy_kernel(const cpx<R>* __restrict__ a, cpx<R>* __restrict__ b, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

__global__ void convert_c128_c64_kernel(const cpx<double>* __restrict__ a, cpx<float>* __restrict__ b, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = cpx<float>((float)a[i].re, (float)a[i].im);
}

__global__ void widen_c64_kernel(const cpx<float>* __restrict__ a, cpx<double>* __restrict__ b, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = cpx<double>((double)a[i].re, (double)a[i].im);
}

// Homogeneous plane waves u_j(r) = d_j exp(-i k d_j . r), row-major
// n x nrhs in the reference's unknown order (node z fastest, then x, then
// y; components u, v, w), r = ((ix-1) dx, (iy-1) dx, (iz-1) dx).
__global__ void planewave_kernel(cpx<double>* __restrict__ x, int nx, int ny, int nz, double dx, double kmag,
                                 const double* __restrict__ dirs, int nrhs)
{
    const int64_t nodes = (int64_t)nx * ny * nz;
    const int64_t total = nodes * nrhs;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(i % nrhs);
        const int64_t g = i / nrhs;
        const int iz = (int)(g % nz), ix = (int)((g / nz) % nx), iy = (int)(g / ((int64_t)nz * nx));
        const double d0 = dirs[3 * j], d1 = dirs[3 * j + 1], d2 = dirs[3 * j + 2];
        const double ph = -kmag * dx * (d0 * ix + d1 * iy + d2 * iz);
        double sn, cs;
        sincos(ph, &sn, &cs);
        cpx<double>* row = x + 3 * g * nrhs + j;
        row[0] = cpx<double>(d0 * cs, d0 * sn);
        row[nrhs] = cpx<double>(d1 * cs, d1 * sn);
        row[2 * nrhs] = cpx<double>(d2 * cs, d2 * sn);
    }
}

// Per-(segment, column) |v|^2 of a row-major block whose rows are cut into
// contiguous segments (the unknowns of the independent systems of a batch):
// block (x, g) sums rows of segment g in its x-th share, thread t column
// t % cw of every (256 / cw)-th row (whole rows per pass, as colnorm_kernel).
__global__ __launch_bounds__(256) void seg_colnorm_kernel(const cpx<double>* __restrict__ v,
                                                          const int64_t* __restrict__ seg_off, int nrhs, int cw,
                                                          double* __restrict__ partial)
{
    __shared__ double s[256];
    const int g = blockIdx.y, c = threadIdx.x % cw, r0t = threadIdx.x / cw, rstep = 256 / cw;
    const int64_t a = seg_off[g], n = seg_off[g + 1] - a;
    const int64_t rows_per_blk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = a + (int64_t)blockIdx.x * rows_per_blk;
    const int64_t r1 = r0 + rows_per_blk < a + n ? r0 + rows_per_blk : a + n;
    double acc = 0.0;
    if (c < nrhs)
        for (int64_t i = r0 + r0t; i < r1; i += rstep) {
            const cpx<double> x = v[i * nrhs + c];
            acc += x.re * x.re + x.im * x.im;
        }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off >= cw; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if ((int)threadIdx.x < nrhs) partial[((int64_t)g * gridDim.x + blockIdx.x) * nrhs + threadIdx.x] = s[threadIdx.x];
}

__global__ void seg_reduce_kernel(const double* __restrict__ partial, int nblk, int nrhs, double* __restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
    if (c >= nrhs) return;
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) acc += partial[((int64_t)g * nblk + b) * nrhs + c];
    out[(int64_t)g * nrhs + c] = acc;
}

// Zeroes the columns c >= ncol[g] of the rows of segment g.
__global__ void seg_mask_kernel(cpx<double>* __restrict__ v, const int64_t* __restrict__ seg_off,
                                const int32_t* __restrict__ ncol, int nseg, int nrhs)
{
    const int64_t n = seg_off[nseg];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nseg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (seg_off[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        for (int c = ncol[lo]; c < nrhs; ++c) v[i * nrhs + c] = cpx<double>(0.0, 0.0);
    }
}

}  // namespace

static unsigned grid1(int64_t n)
{
    int64_t b = (n + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

template <typename Rs, typename R>
int wt_to_elim(const cpx<Rs>* src, const int64_t* iperm, const double* scale, cpx<R>* y, int64_t n, int nrhs,
               hipStream_t s)
{
    dim3 grid((unsigned)((n + TP - 1) / TP), (nrhs + TP - 1) / TP);
    hipLaunchKernelGGL((to_elim_kernel<Rs, R>), grid, dim3(256), 0, s, src, iperm, scale, y, n, nrhs);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template <typename R, typename Rd>
int wt_from_elim(const cpx<R>* y, const int64_t* iperm, cpx<Rd>* dst, int64_t n, int nrhs, int accumulate,
                 hipStream_t s)
{
    dim3 grid((unsigned)((n + TP - 1) / TP), (nrhs + TP - 1) / TP);
    hipLaunchKernelGGL((from_elim_kernel<R, Rd>), grid, dim3(256), 0, s, y, iperm, dst, n, nrhs, accumulate);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

int64_t wt_residual_blocks(int64_t n) { return (n + SP_ROWS - 1) / SP_ROWS; }

template <typename Ra>
int wt_residual(const int32_t* rowptr, const int32_t* colidx, const cpx<Ra>* vals, const cpx<double>* x,
                const cpx<double>* b, cpx<double>* r, double* partial, double* rnorm2, int64_t n, int nrhs,
                hipStream_t s)
{
    if (nrhs < 1 || nrhs > 64) return WT_ERR_ARG;  // one wave per row covers up to 64 columns
    if (nrhs > 16 && n % 3 == 0) {
        const int64_t nb = (n / 3 + 3) / 4;  // <= wt_residual_blocks(n)
        if (b)
            hipLaunchKernelGGL((residual_node_kernel<Ra, float, true, RES_TB>), dim3((unsigned)nb), dim3(256), 0, s,
                               rowptr, colidx, vals, x, b, r, nullptr, nullptr, nullptr, partial, n / 3, nrhs);
        else
            hipLaunchKernelGGL((residual_node_kernel<Ra, float, false, RES_TB>), dim3((unsigned)nb), dim3(256), 0, s,
                               rowptr, colidx, vals, x, b, r, nullptr, nullptr, nullptr, partial, n / 3, nrhs);
        WT_HIP(hipGetLastError());
        hipLaunchKernelGGL(reduce_cols_kernel, dim3(nrhs), dim3(256), 0, s, partial, nb, nrhs, rnorm2);
        WT_HIP(hipGetLastError());
        return WT_OK;
    }
    int cw = 1;
    while (cw < nrhs) cw <<= 1;
    const int64_t rows_per_blk = SP_ROWS * (64 / cw);
    const int64_t nb = (n + rows_per_blk - 1) / rows_per_blk;  // <= wt_residual_blocks(n)
    switch (cw) {
#define WT_RES(CW)                                                                                              \
    case CW:                                                                                                    \
        if (b)                                                                                                  \
            hipLaunchKernelGGL((residual_kernel<Ra, CW, true>), dim3((unsigned)nb), dim3(256), 0, s, rowptr,    \
                               colidx, vals, x, b, r, partial, n, nrhs);                                        \
        else                                                                                                    \
            hipLaunchKernelGGL((residual_kernel<Ra, CW, false>), dim3((unsigned)nb), dim3(256), 0, s, rowptr,   \
                               colidx, vals, x, b, r, partial, n, nrhs);                                        \
        break;
        WT_RES(1) WT_RES(2) WT_RES(4) WT_RES(8) WT_RES(16) WT_RES(32) WT_RES(64)
#undef WT_RES
    }
    WT_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_cols_kernel, dim3(nrhs), dim3(256), 0, s, partial, nb, nrhs, rnorm2);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

bool wt_residual_elim_ok(int64_t n, int nrhs) { return nrhs > 16 && nrhs <= 64 && n % 3 == 0; }

template <typename Ra, typename Ry>
int wt_residual_elim(const int32_t* rowptr, const int32_t* colidx, const cpx<Ra>* vals, const cpx<double>* x,
                     const cpx<double>* b, const int64_t* perm, const double* scale, cpx<Ry>* y, double* partial,
                     double* rnorm2, int64_t n, int nrhs, hipStream_t s)
{
    if (!wt_residual_elim_ok(n, nrhs) || !perm || !y || !b) return WT_ERR_ARG;
    const int64_t nb = (n / 3 + 3) / 4;  // <= wt_residual_blocks(n)
    hipLaunchKernelGGL((residual_node_kernel<Ra, Ry, true, RES_TB>), dim3((unsigned)nb), dim3(256), 0, s, rowptr,
                       colidx, vals, x, b, nullptr, perm, scale, y, partial, n / 3, nrhs);
    WT_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_cols_kernel, dim3(nrhs), dim3(256), 0, s, partial, nb, nrhs, rnorm2);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

int wt_colnorm2(const cpx<double>* v, int64_t n, int nrhs, double* partial, int64_t nblk, double* out, hipStream_t s)
{
    if (nrhs < 1 || nrhs > 256) return WT_ERR_ARG;
    int cw = 1;
    while (cw < nrhs) cw <<= 1;
    hipLaunchKernelGGL(colnorm_kernel, dim3((unsigned)nblk), dim3(256), 0, s, v, n, nrhs, cw, partial);
    WT_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_cols_kernel, dim3(nrhs), dim3(256), 0, s, partial, nblk, nrhs, out);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template <typename R>
int wt_copy(const cpx<R>* a, cpx<R>* b, int64_t n, hipStream_t s)
{
    if (n <= 0) return WT_OK;
    hipLaunchKernelGGL(copy_kernel<R>, dim3(grid1(n)), dim3(256), 0, s, a, b, n);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

int wt_convert_c128_c64(const cpx<double>* a, cpx<float>* b, int64_t n, hipStream_t s)
{
    if (n <= 0) return WT_OK;
    hipLaunchKernelGGL(convert_c128_c64_kernel, dim3(grid1(n)), dim3(256), 0, s, a, b, n);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

int wt_planewave(cpx<double>* x, int nx, int ny, int nz, double dx, double kmag, const double* d_dirs, int nrhs,
                 hipStream_t s)
{
    const int64_t total = (int64_t)nx * ny * nz * nrhs;
    hipLaunchKernelGGL(planewave_kernel, dim3(grid1(total)), dim3(256), 0, s, x, nx, ny, nz, dx, kmag, d_dirs, nrhs);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

int wt_widen_c64(const cpx<float>* a, cpx<double>* b, int64_t n, hipStream_t s)
{
    if (n <= 0) return WT_OK;
    hipLaunchKernelGGL(widen_c64_kernel, dim3(grid1(n)), dim3(256), 0, s, a, b, n);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template int wt_to_elim<double, float>(const cpx<double>*, const int64_t*, const double*, cpx<float>*, int64_t, int,
                                       hipStream_t);
template int wt_to_elim<double, double>(const cpx<double>*, const int64_t*, const double*, cpx<double>*, int64_t,
                                        int, hipStream_t);
template int wt_from_elim<float, double>(const cpx<float>*, const int64_t*, cpx<double>*, int64_t, int, int,
                                         hipStream_t);
template int wt_from_elim<double, double>(const cpx<double>*, const int64_t*, cpx<double>*, int64_t, int, int,
                                          hipStream_t);
template int wt_residual<double>(const int32_t*, const int32_t*, const cpx<double>*, const cpx<double>*,
                                 const cpx<double>*, cpx<double>*, double*, double*, int64_t, int, hipStream_t);
template int wt_residual<float>(const int32_t*, const int32_t*, const cpx<float>*, const cpx<double>*,
                                const cpx<double>*, cpx<double>*, double*, double*, int64_t, int, hipStream_t);
template int wt_residual_elim<double, float>(const int32_t*, const int32_t*, const cpx<double>*, const cpx<double>*,
                                             const cpx<double>*, const int64_t*, const double*, cpx<float>*, double*,
                                             double*, int64_t, int, hipStream_t);
template int wt_residual_elim<double, double>(const int32_t*, const int32_t*, const cpx<double>*, const cpx<double>*,
                                              const cpx<double>*, const int64_t*, const double*, cpx<double>*,
                                              double*, double*, int64_t, int, hipStream_t);
template int wt_copy<double>(const cpx<double>*, cpx<double>*, int64_t, hipStream_t);
template int wt_copy<float>(const cpx<float>*, cpx<float>*, int64_t, hipStream_t);

int wt_seg_colnorm2(const cpx<double>* v, const int64_t* d_seg_off, int nseg, int nrhs, double* partial, int nblk,
                    double* out, hipStream_t s)
{
    if (nseg < 1 || nrhs < 1 || nblk < 1 || nseg > 65535 || nrhs > 256) return WT_ERR_ARG;
    int cw = 1;
    while (cw < nrhs) cw <<= 1;
    hipLaunchKernelGGL(seg_colnorm_kernel, dim3((unsigned)nblk, nseg), dim3(256), 0, s, v, d_seg_off, nrhs, cw,
                       partial);
    WT_HIP(hipGetLastError());
    hipLaunchKernelGGL(seg_reduce_kernel, dim3((nrhs + 63) / 64, nseg), dim3(64), 0, s, partial, nblk, nrhs, out);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

int wt_seg_mask_cols(cpx<double>* v, const int64_t* d_seg_off, const int32_t* d_ncol, int nseg, int64_t n, int nrhs,
                     hipStream_t s)
{
    if (nseg < 1 || n <= 0) return WT_OK;
    hipLaunchKernelGGL(seg_mask_kernel, dim3(grid1(n)), dim3(256), 0, s, v, d_seg_off, d_ncol, nseg, nrhs);
    WT_HIP(hipGetLastError());
    return WT_OK;
}
