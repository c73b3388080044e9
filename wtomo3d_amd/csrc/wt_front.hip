// Front-level kernels of the multifrontal factorization and solve.
//
// Replaces the numeric phase STRUMPACK runs in job 2 (strumpack_finter.c:
// 68-92, STRUMPACK_factor) and the dense parts of job 3 (STRUMPACK_solve).
// Each front stores [L11\U11; L21] as a (p+u) x p column block and U12 as a
// p x u block; the u x u contribution block (CB) lives in an arena indexed by
// tree level parity.  Partial pivoting is done inside each nb x nb diagonal
// tile; the tile's unit-lower and upper inverses are kept so the panel
// solves and the forward/back substitutions run as MFMA GEMMs.
#include "wt_front.h"

#ifndef FILL16
#define FILL16 1
#endif

namespace {

template <typename R>
__global__ void fill_zero_kernel(cpx<R>* __restrict__ p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = cpx<R>(R(0), R(0));
}

// Zeroes n16 16-byte units, four stores per thread and pass (consecutive
// lanes on consecutive units in each store).
__global__ __launch_bounds__(256) void fill_zero16_kernel(uint4* __restrict__ p, int64_t n16)
{
    const int64_t nt = (int64_t)gridDim.x * blockDim.x;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += 4 * nt) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k * nt < n16) p[i + k * nt] = z;
    }
}

template <typename Ri, typename R>
__global__ void scatter_kernel(const cpx<Ri>* __restrict__ vals, const int64_t* __restrict__ dst, int64_t nnz,
                               cpx<R>* __restrict__ arena)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < nnz; i += stride) {
        cpx<Ri> v = vals[i];
        arena[dst[i]] = cpx<R>((R)v.re, (R)v.im);
    }
}

// Row equilibration: scale[i] = 2^-e with 2^e <= max_j max(|re a_ij|,
// |im a_ij|) < 2^(e+1), an exact power of two, so the scaled system D A x =
// D b carries no extra rounding.  One wave per row.
template <typename Ri>
__global__ __launch_bounds__(256) void row_scale_kernel(const int32_t* __restrict__ rowptr,
                                                        const cpx<Ri>* __restrict__ vals, int64_t n,
                                                        double* __restrict__ scale)
{
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    double m = 0.0;
    for (int32_t e = rowptr[row] + lane; e < rowptr[row + 1]; e += 64) {
        double a = fabs((double)vals[e].re), b = fabs((double)vals[e].im);
        m = fmax(m, fmax(a, b));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if (lane == 0) {
        int ex = 0;
        (void)frexp(m, &ex);  // m = f 2^ex, f in [0.5, 1)
        scale[row] = m > 0.0 ? ldexp(1.0, 1 - ex) : 1.0;
    }
}

// arena[dst[e]] = scale[row] * vals[e] for every entry of the CSR rows.
template <typename Ri, typename R>
__global__ __launch_bounds__(256) void scatter_rows_kernel(const int32_t* __restrict__ rowptr,
                                                           const cpx<Ri>* __restrict__ vals,
                                                           const int64_t* __restrict__ dst,
                                                           const double* __restrict__ scale, int64_t n,
                                                           cpx<R>* __restrict__ arena)
{
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double sc = scale[row];
    for (int32_t e = rowptr[row] + lane; e < rowptr[row + 1]; e += 64) {
        cpx<Ri> v = vals[e];
        arena[dst[e]] = cpx<R>((R)(sc * (double)v.re), (R)(sc * (double)v.im));
    }
}


#ifndef EA_UNROLL
#define EA_UNROLL 4
#endif

// Adds the contribution blocks of the listed children into their parents.
// At most one child per parent is listed, so the adds do not race.  The
// child-to-parent map ascends, so the child's indices below q (the first
// one past the parent's pivots) land on the parent's pivot rows/columns:
// part 1 adds every entry with a row or column below q (into the parent's
// L and U blocks), part 2 the rest (into the parent's contribution block),
// part 0 everything.
template <typename R>
__global__ __launch_bounds__(256) void extend_add_kernel(const wt_dfront* __restrict__ fr,
                                                         const int32_t* __restrict__ kids,
                                                         const int32_t* __restrict__ col_start, int nkids,
                                                         const int32_t* __restrict__ relmap,
                                                         const cpx<R>* __restrict__ cb_child,
                                                         cpx<R>* __restrict__ factor,
                                                         cpx<R>* __restrict__ cb_parent, int part)
{
    // child of this block: the last i with col_start[i] <= block
    const int b = (int)blockIdx.x;
    int lo = 0, hi = nkids - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (col_start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const wt_dfront C = fr[kids[lo]];
    const int64_t uc = C.u;
    const int64_t j0 = (int64_t)(b - col_start[lo]) * WT_EA_COLS;
    if (j0 >= uc) return;
    const wt_dfront P = fr[C.parent];
    const int32_t* map = relmap + C.moff;
    const int64_t pp = P.p, ld = P.p + P.u, pu = P.u;
    int64_t q = 0;
    if (part) {
        int64_t hi = uc;
        while (q < hi) {
            const int64_t mid = (q + hi) >> 1;
            if (map[mid] < pp) q = mid + 1;
            else hi = mid;
        }
    }
    const int64_t j1 = j0 + WT_EA_COLS < uc ? j0 + WT_EA_COLS : uc;
    if (part == 2 && j1 <= q) return;
    const cpx<R>* cb = cb_child + C.CBoff;
    cpx<R>* Lp = factor + P.Loff;
    cpx<R>* Up = factor + P.Uoff;
    cpx<R>* Cp = cb_parent + P.CBoff;
    // EA_UNROLL entries per thread and pass: their index, child and target
    // loads are all issued before the first store.  The map is injective, so
    // the targets of one pass are distinct and every entry still receives
    // exactly one add.
    constexpr int UR = EA_UNROLL;
    const int64_t nt = blockDim.x;
    for (int64_t j = (part == 2 && j0 < q) ? q : j0; j < j1; ++j) {
        const int64_t s = map[j];
        const int64_t i0 = part == 2 ? q : 0, i1 = (part == 1 && j >= q) ? q : uc;
        const cpx<R>* cbj = cb + j * uc;
        for (int64_t ib = i0 + threadIdx.x; ib < i1; ib += UR * nt) {
            int64_t r[UR];
            cpx<R> v[UR], o[UR];
            cpx<R>* t[UR];
#pragma unroll
            for (int k = 0; k < UR; ++k) {
                const int64_t i = ib + k * nt;
                r[k] = i < i1 ? map[i] : 0;
                v[k] = i < i1 ? cbj[i] : cpx<R>(0, 0);
            }
#pragma unroll
            for (int k = 0; k < UR; ++k) {
                if (s < pp) t[k] = Lp + s * ld + r[k];
                else if (r[k] < pp) t[k] = Up + (s - pp) * pp + r[k];
                else t[k] = Cp + (s - pp) * pu + (r[k] - pp);
                if (ib + k * nt < i1) o[k] = *t[k];
            }
#pragma unroll
            for (int k = 0; k < UR; ++k)
                if (ib + k * nt < i1) *t[k] = o[k] + v[k];
        }
    }
}

// LU with partial pivoting of the 64 x 64 diagonal tile starting at pivot kb
// of every listed front, plus the inverses of its unit-lower and upper
// factors.  perm[i] = tile row that ends at position i.
//
// One 256-thread block per tile.  The four waves load and store the tile
// together (every load of a thread in flight at once); the elimination runs
// with lane = row, wave 0 choosing the pivot (largest |re|+|im| in the
// column, first on ties) and swapping whole tile rows, and the columns of
// each rank-1 update split over the waves.  The inverses are swept row by
// row with lane = column, so the factor entry of every step is a
// wave-uniform LDS broadcast: wave 0 builds L^{-1} while wave 1 builds
// U^{-1} (one after the other in FP64, whose second buffer would not fit in
// LDS) and the other waves write the factored tile out.
// Adds one tile's pivot statistics to the factorization's (see
// WT_PIVSTAT_LEN): one call per tile, by one lane.
__device__ __forceinline__ void pivstat_publish(int* status, int ntiny, float min_ratio, float growth)
{
    if (ntiny) atomicAdd(status + 1, ntiny);
    atomicMin((unsigned*)status + 2, __float_as_uint(min_ratio));
    atomicMax((unsigned*)status + 3, __float_as_uint(growth));
}

// The tiny-pivot threshold relative to the tile's largest entry: about the
// square root of the precision's unit roundoff (2^-12 for fp32, 2^-26 for
// fp64).
template <typename R>
__device__ constexpr R tiny_pivot_rel()
{
    return sizeof(R) == 4 ? R(2.44140625e-4) : R(1.490116119384765625e-8);
}

template <typename R>
__global__ __launch_bounds__(256) void diag_tile_kernel(const wt_dfront* __restrict__ fr,
                                                        const int32_t* __restrict__ list, int64_t kb,
                                                        cpx<R>* __restrict__ factor, cpx<R>* __restrict__ tiles,
                                                        int32_t* __restrict__ perm, int* __restrict__ status)
{
    constexpr int NB = WT_NB, CPT = NB / 4;
    constexpr int NX = sizeof(R) == 4 ? 2 : 1;  // inverse buffers that fit beside T
    __shared__ cpx<R> T[NB][NB + 1];          // T[col][row]
    __shared__ cpx<R> X[NX][NB][NB + 1];      // X[.][row][col] of an inverse
    __shared__ R wmax[4];                     // per wave: max |re|+|im| of the loaded tile
    const wt_dfront F = fr[list[blockIdx.x]];
    if (kb >= F.p) return;
    const int w = (int)(F.p - kb < NB ? F.p - kb : NB);
    const int64_t ld = F.p + F.u;
    cpx<R>* base = factor + F.Loff + kb * ld + kb;
    const int tid = threadIdx.x, lane = tid & 63, wq = tid >> 6;
    {
        cpx<R> v[CPT];
        R amax = R(0);
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int c = wq + 4 * k;
            v[k] = (c < w && lane < w) ? base[(int64_t)c * ld + lane] : cpx<R>(R(0), R(0));
            amax = fmax(amax, fabs(v[k].re) + fabs(v[k].im));
        }
#pragma unroll
        for (int k = 0; k < CPT; ++k) T[wq + 4 * k][lane] = v[k];
        for (int off = 32; off > 0; off >>= 1) amax = fmax(amax, __shfl_xor(amax, off));
        if (lane == 0) wmax[wq] = amax;
    }
    int myperm = lane;  // wave 0: tile row that ends at position `lane`
    __syncthreads();
    const R a0 = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    float min_ratio = __builtin_inff();  // thread 0's running pivot statistics
    int ntiny = 0;
    for (int j = 0; j < w; ++j) {
        if (wq == 0) {
            R a = R(-1);
            if (lane >= j && lane < w) a = fabs(T[j][lane].re) + fabs(T[j][lane].im);
            int bi = lane;
            for (int off = 32; off > 0; off >>= 1) {
                const R oa = __shfl_xor(a, off);
                const int oi = __shfl_xor(bi, off);
                if (oa > a || (oa == a && oi < bi)) { a = oa; bi = oi; }
            }
            const int pv = bi;  // uniform across the wave
            if (pv != j) {
                if (lane < w) {
                    const cpx<R> t0 = T[lane][j];
                    T[lane][j] = T[lane][pv];
                    T[lane][pv] = t0;
                }
                const int pj = __shfl(myperm, j), pp = __shfl(myperm, pv);
                if (lane == j) myperm = pp;
                if (lane == pv) myperm = pj;
            }
        }
        __syncthreads();
        cpx<R> d = T[j][j];
        if (tid == 0) {
            const R ad = fabs(d.re) + fabs(d.im);
            min_ratio = fminf(min_ratio, a0 > R(0) ? (float)(ad / a0) : 0.f);
            ntiny += ad <= tiny_pivot_rel<R>() * a0;
        }
        if (d.re == R(0) && d.im == R(0)) {
            if (tid == 0) atomicOr(status, 1);
            d = cpx<R>(R(1e-30), R(0));
        }
        const cpx<R> dinv = cdiv(cpx<R>(R(1), R(0)), d);
        cpx<R> l(R(0), R(0));
        if (lane > j && lane < w) {
            l = T[j][lane] * dinv;
            for (int c = j + 1 + wq; c < w; c += 4) T[c][lane] = T[c][lane] - l * T[c][j];
        }
        __syncthreads();
        // column j is read again only by wave 0 (next swap), after these writes
        if (wq == 0) {
            if (lane == j) T[j][j] = d;
            if (lane > j && lane < w) T[j][lane] = l;
        }
    }
    __syncthreads();
    if (wq == 0 && lane < w) perm[F.pivoff + kb + lane] = myperm;
    cpx<R>* Li = tiles + F.tileoff + (kb / NB) * 2 * NB * NB;
    cpx<R>* Ui = Li + NB * NB;
    const int c = lane;
    // L^{-1}[i][c] = delta_ic - sum_{k<i} L[i][k] L^{-1}[k][c]   (zero above the diagonal)
    auto lower_inv = [&](cpx<R> (*XL)[NB + 1]) {
        for (int i = 0; i < w; ++i) {
            cpx<R> s0(R(0), R(0)), s1(R(0), R(0));
            int k = 0;
            for (; k + 1 < i; k += 2) {
                s0 = s0 + T[k][i] * XL[k][c];
                s1 = s1 + T[k + 1][i] * XL[k + 1][c];
            }
            if (k < i) s0 = s0 + T[k][i] * XL[k][c];
            XL[i][c] = i == c ? cpx<R>(R(1), R(0)) : (i < c ? cpx<R>(R(0), R(0)) : -(s0 + s1));
        }
    };
    // U^{-1}[i][c] = (delta_ic - sum_{k>i} U[i][k] U^{-1}[k][c]) / U[i][i]   (zero below)
    auto upper_inv = [&](cpx<R> (*XU)[NB + 1]) {
        for (int i = w - 1; i >= 0; --i) {
            cpx<R> s0(R(0), R(0)), s1(R(0), R(0));
            int k = i + 1;
            for (; k + 1 < w; k += 2) {
                s0 = s0 + T[k][i] * XU[k][c];
                s1 = s1 + T[k + 1][i] * XU[k + 1][c];
            }
            if (k < w) s0 = s0 + T[k][i] * XU[k][c];
            const cpx<R> rdi = cdiv(cpx<R>(R(1), R(0)), T[i][i]);
            XU[i][c] = i > c ? cpx<R>(R(0), R(0)) : (i == c ? rdi : -((s0 + s1) * rdi));
        }
    };
    auto store_inv = [&](cpx<R> (*Xs)[NB + 1], cpx<R>* dst) {
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int cc = wq + 4 * k;
            if (cc < w && lane < w) dst[(int64_t)cc * NB + lane] = Xs[lane][cc];
        }
    };
    if (NX == 2) {
        if (wq == 0) lower_inv(X[0]);
        else if (wq == 1) upper_inv(X[NX - 1]);
    } else if (wq == 0) {
        lower_inv(X[0]);
    }
    R umax = R(0);  // largest |U| entry among this thread's stores
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int cc = wq + 4 * k;
        if (cc < w && lane < w) {
            const cpx<R> t = T[cc][lane];
            base[(int64_t)cc * ld + lane] = t;
            if (lane <= cc) umax = fmax(umax, fabs(t.re) + fabs(t.im));
        }
    }
    for (int off = 32; off > 0; off >>= 1) umax = fmax(umax, __shfl_xor(umax, off));
    if (lane == 0) {
        const float growth = a0 > R(0) ? (float)(umax / a0) : 0.f;
        if (wq == 0) pivstat_publish(status, ntiny, min_ratio, growth);
        else atomicMax((unsigned*)status + 3, __float_as_uint(growth));
    }
    __syncthreads();
    store_inv(X[0], Li);
    if (NX == 1) {
        __syncthreads();
        if (wq == 1) upper_inv(X[0]);
        __syncthreads();
    }
    store_inv(X[NX - 1], Ui);
}

// acc += a x and acc -= a x as chains of fused multiply-adds, so that no
// product is held apart from its accumulator
__device__ __forceinline__ void cmac(cpx<float>& acc, cpx<float> a, cpx<float> x)
{
    acc.re = fmaf(a.re, x.re, acc.re);
    acc.re = fmaf(-a.im, x.im, acc.re);
    acc.im = fmaf(a.re, x.im, acc.im);
    acc.im = fmaf(a.im, x.re, acc.im);
}
__device__ __forceinline__ void cmsub(cpx<float>& acc, cpx<float> a, cpx<float> x)
{
    acc.re = fmaf(-a.re, x.re, acc.re);
    acc.re = fmaf(a.im, x.im, acc.re);
    acc.im = fmaf(-a.re, x.im, acc.im);
    acc.im = fmaf(-a.im, x.re, acc.im);
}

// The same tile LU and inverses in FP32 with the tile in registers: one
// 128-thread block per tile.  Wave 0 holds row `lane` of the tile in 64
// complex registers and eliminates without moving rows: at step j the pivot
// is the not-yet-chosen row with the largest |re|+|im| in column j (the
// lowest row on ties), its row is read across lanes with v_readlane, and the
// other open rows take their multiplier and update in place.  The row chosen
// at step j becomes position j, so every lane ends up holding its own row of
// the packed L\U and stores it straight to its position; there are no block
// barriers inside the elimination.  Both inverses then run as the same
// unit-lower forward substitution with lane = column, fully unrolled over
// registers with the matrix entries as LDS broadcasts: wave 0 inverts L,
// wave 1 inverts U~^T where U = D U~ (D the pivots), which yields the rows of
// U~^{-1} and so U^{-1} = U~^{-1} D^{-1} with coalesced stores; L^{-1} goes out
// through an LDS transpose.  Entries outside the w x w tile are stored as 0.
__device__ __forceinline__ float wave_max(float v)
{
    // max within rows of 16 by DPP, then across the four rows
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, false)));
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

__device__ __forceinline__ cpx<float> read_lane(cpx<float> v, int l)
{
    return cpx<float>(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.re), l)),
                      __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.im), l)));
}

__global__ __launch_bounds__(128) void diag_tile_wave_kernel(const wt_dfront* __restrict__ fr,
                                                             const int32_t* __restrict__ list, int64_t kb,
                                                             cpx<float>* __restrict__ factor,
                                                             cpx<float>* __restrict__ tiles,
                                                             int32_t* __restrict__ perm, int* __restrict__ status)
{
    typedef cpx<float> C;
    constexpr int NB = WT_NB;
    __shared__ C M[2][NB][NB + 1];  // [0]: L[i][k] at [k][i], later L^{-1} [i][c]; [1]: U~^T[i][k] at [k][i]
    __shared__ float Dre[NB], Dim[NB];  // 1 / pivot
    const wt_dfront F = fr[list[blockIdx.x]];
    if (kb >= F.p) return;
    const int w = (int)(F.p - kb < NB ? F.p - kb : NB);
    const int64_t ld = F.p + F.u;
    C* base = factor + F.Loff + kb * ld + kb;
    const int lane = threadIdx.x & 63, wq = threadIdx.x >> 6;
    if (wq == 0) {
        C a[NB];
        {
            const C* src = base + (lane < w ? lane : 0);
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                const C v = src[(int64_t)(c < w ? c : w - 1) * ld];
                a[c] = (lane < w && c < w) ? v : C(0.f, 0.f);
            }
        }
        bool open = lane < w;  // not yet chosen as a pivot row
        int pos = lane;        // position of this row once chosen
        C dme(1.f, 0.f);       // 1 / own pivot
        float a0 = 0.f;        // max |re|+|im| of the tile (entries outside it are 0)
#pragma unroll
        for (int c = 0; c < NB; ++c) a0 = fmaxf(a0, fabsf(a[c].re) + fabsf(a[c].im));
        a0 = wave_max(a0);
        float pab = 0.f;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < w) {
                const float m = open ? fabsf(a[j].re) + fabsf(a[j].im) : -1.f;
                const float mx = wave_max(m);
                uint64_t hit = __ballot(open && m == mx);
                if (hit == 0) hit = __ballot(open);  // NaN column: first open row
                const int pv = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(hit));
                C d = read_lane(a[j], pv);
                // lane j keeps |pivot j| for the statistics taken after the
                // loop (kept out of this fully unrolled loop, which holds
                // the tile in registers only while its body stays small)
                if (lane == j) pab = fabsf(d.re) + fabsf(d.im);
                if (d.re == 0.f && d.im == 0.f) {
                    if (lane == 0) atomicOr(status, 1);
                    d = C(1e-30f, 0.f);
                }
                const C dinv = cdiv(C(1.f, 0.f), d);
                const bool elim = open && lane != pv;
                if (lane == pv) {
                    open = false;
                    pos = j;
                    dme = dinv;
                }
                if (lane == 0) { Dre[j] = dinv.re; Dim[j] = dinv.im; }
                // rows not eliminated here take l = 0 and keep their values
                const C l = elim ? a[j] * dinv : C(0.f, 0.f);
                if (elim) a[j] = l;
#pragma unroll
                for (int c = j + 1; c < NB; ++c) cmsub(a[c], l, read_lane(a[c], pv));
            } else if (lane == 0) {
                Dre[j] = 0.f;
                Dim[j] = 0.f;
            }
        }
        if (lane < w) perm[F.pivoff + kb + pos] = lane;
        {
            // this row is row `pos` of U from column pos on
            float umax = 0.f;
#pragma unroll
            for (int c = 0; c < NB; ++c)
                if (c >= pos) umax = fmaxf(umax, fabsf(a[c].re) + fabsf(a[c].im));
            umax = wave_max(lane < w ? umax : 0.f);
            float ratio = lane < w ? (a0 > 0.f ? pab / a0 : 0.f) : __builtin_inff();
            const int ntiny = __popcll(__ballot(lane < w && pab <= tiny_pivot_rel<float>() * a0));
            for (int off = 32; off > 0; off >>= 1) ratio = fminf(ratio, __shfl_xor(ratio, off));
            if (lane == 0) pivstat_publish(status, ntiny, ratio, a0 > 0.f ? umax / a0 : 0.f);
        }
        // packed L\U to its positions; L below the diagonal and U~^T = (D^{-1} U)^T to LDS
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            if (lane < w && c < w) base[(int64_t)c * ld + pos] = a[c];
            M[0][c][pos] = c < pos ? a[c] : C(0.f, 0.f);
            M[1][pos][c] = c > pos ? dme * a[c] : C(0.f, 0.f);
        }
    }
    __syncthreads();
    // x = column `lane` of the inverse of the unit-lower M[wq]
    const C (*Mw)[NB + 1] = M[wq];
    C x[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) x[i] = C(i == lane ? 1.f : 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < NB - 1; ++k) {
#pragma unroll
        for (int i = k + 1; i < NB; ++i) cmsub(x[i], Mw[k][i], x[k]);
    }
    C* dst = tiles + F.tileoff + (kb / NB) * 2 * NB * NB;
    if (wq == 1) {
        // x[i] = U~^{-1}[lane][i]; U^{-1}[lane][i] = x[i] / pivot_i, column-major at i * NB + lane
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const C v = x[i] * C(Dre[i], Dim[i]);
            dst[NB * NB + i * NB + lane] = (lane < w && i < w) ? v : C(0.f, 0.f);
        }
    }
    __syncthreads();  // wave 1 is done with M[1]; wave 0 has read all of M[0]
    if (wq == 0) {
#pragma unroll
        for (int i = 0; i < NB; ++i) M[0][i][lane] = x[i];
    }
    __syncthreads();
    // L^{-1} column-major: both waves store half of the columns
#pragma unroll 4
    for (int cc = wq; cc < NB; cc += 2)
        dst[cc * NB + lane] = (lane < w && cc < w) ? M[0][lane][cc] : C(0.f, 0.f);
}

// ---- super-panel triangular solves of the block solve ----
//
// One launch covers one super-panel (<= WT_NB2 pivots) of every listed front
// (entries {front, K0, K1}); blocks split the right-hand sides into groups
// of CG columns, which are independent, so a block walks the panel's
// 64-pivot tiles sequentially with no inter-block synchronization.  Wave q
// of the block's four owns columns [q CG/4, (q+1) CG/4) of the group and
// lane l owns rows l, l+64, ...: every factor entry a lane loads feeds CG/4
// complex multiply-adds, and the right-hand-side values come from LDS as
// wave-uniform broadcasts.
// acc[c] += sum_k A[k*lda + row] * V[k][cb + c] over k < w (V zero past w).
// Factor entries are loaded in batches of KB, all in flight together: the
// launches of the top of the tree have few blocks and are latency bound, so
// their 4-column groups take the whole 64-entry row at once; the 16-column
// groups of the wide launches keep a smaller batch for occupancy.
template <typename R, int CG, int CPW>
__device__ __forceinline__ void sp_dot(const cpx<R>* __restrict__ A, int64_t lda, int w,
                                       const cpx<R> (*V)[CG], int cb, cpx<R>* acc)
{
    constexpr int KB = CG == 4 ? (sizeof(R) == 4 ? 64 : 32) : 16;
    for (int k0 = 0; k0 < w; k0 += KB) {
        cpx<R> a[KB];
#pragma unroll
        for (int j = 0; j < KB; ++j)
            a[j] = k0 + j < w ? A[(int64_t)(k0 + j) * lda] : cpx<R>(R(0), R(0));
#pragma unroll
        for (int j = 0; j < KB; ++j)
#pragma unroll
            for (int c = 0; c < CPW; ++c) acc[c] = acc[c] + a[j] * V[k0 + j][cb + c];
    }
}

// forward: y_t = L_t^{-1} P_t y_t, then y[t+w..K1) -= L[t+w..K1, t] y_t
template <typename R, int CG>
__global__ __launch_bounds__(256) void sp_fwd_kernel(const wt_dfront* __restrict__ fr, const int32_t* __restrict__ ent,
                                                     const cpx<R>* __restrict__ factor,
                                                     const cpx<R>* __restrict__ tiles,
                                                     const int32_t* __restrict__ perm, cpx<R>* __restrict__ y,
                                                     int64_t n, int nrhs)
{
    constexpr int NB = WT_NB, CPW = CG / 4;
    __shared__ cpx<R> yt[NB][CG];
    __shared__ cpx<R> xt[NB][CG];
    const int32_t* e = ent + 3 * blockIdx.x;
    const wt_dfront F = fr[e[0]];
    const int64_t K0 = e[1], K1 = e[2];
    const int c0 = blockIdx.y * CG;
    if (c0 >= nrhs) return;
    const int nc = nrhs - c0 < CG ? nrhs - c0 : CG;
    const int64_t ld = F.p + F.u;
    const cpx<R>* L = factor + F.Loff;
    cpx<R>* yb = y + F.pstart + (int64_t)c0 * n;
    const int tid = threadIdx.x, lane = tid & 63, cb = (tid >> 6) * CPW;
    for (int64_t t = K0; t < K1; t += NB) {
        const int w = (int)(K1 - t < NB ? K1 - t : NB);
        const int32_t* pr = perm + F.pivoff + t;
        for (int idx = tid; idx < NB * CG; idx += blockDim.x) {
            const int i = idx % NB, c = idx / NB;
            yt[i][c] = (c < nc && i < w) ? yb[(int64_t)c * n + t + pr[i]] : cpx<R>(R(0), R(0));
        }
        __syncthreads();
        // the inverse holds explicit zeros above its diagonal inside the
        // tile; entries past the tile width are never read
        if (lane < w) {
            const cpx<R>* Li = tiles + F.tileoff + (t / NB) * 2 * NB * NB + lane;
            cpx<R> acc[CPW];
#pragma unroll
            for (int c = 0; c < CPW; ++c) acc[c] = cpx<R>(R(0), R(0));
            sp_dot<R, CG, CPW>(Li, NB, w, yt, cb, acc);
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                xt[lane][cb + c] = acc[c];
                if (cb + c < nc) yb[(int64_t)(cb + c) * n + t + lane] = acc[c];
            }
        } else {
#pragma unroll
            for (int c = 0; c < CPW; ++c) xt[lane][cb + c] = cpx<R>(R(0), R(0));
        }
        __syncthreads();
        const int rows = (int)(K1 - t - w);
        for (int r = lane; r < rows; r += 64) {
            const int64_t row = t + w + r;
            cpx<R> acc[CPW];
#pragma unroll
            for (int c = 0; c < CPW; ++c) acc[c] = cpx<R>(R(0), R(0));
            sp_dot<R, CG, CPW>(L + t * ld + row, ld, w, xt, cb, acc);
#pragma unroll
            for (int c = 0; c < CPW; ++c)
                if (cb + c < nc) {
                    cpx<R>* q = yb + (int64_t)(cb + c) * n + row;
                    *q = *q - acc[c];
                }
        }
        __syncthreads();
    }
}

// backward: x_t = U_t^{-1} y_t, then y[K0..t) -= U[K0..t, t] x_t, tiles descending
template <typename R, int CG>
__global__ __launch_bounds__(256) void sp_bwd_kernel(const wt_dfront* __restrict__ fr, const int32_t* __restrict__ ent,
                                                     const cpx<R>* __restrict__ factor,
                                                     const cpx<R>* __restrict__ tiles, cpx<R>* __restrict__ y,
                                                     int64_t n, int nrhs)
{
    constexpr int NB = WT_NB, CPW = CG / 4;
    __shared__ cpx<R> yt[NB][CG];
    __shared__ cpx<R> xt[NB][CG];
    const int32_t* e = ent + 3 * blockIdx.x;
    const wt_dfront F = fr[e[0]];
    const int64_t K0 = e[1], K1 = e[2];
    const int c0 = blockIdx.y * CG;
    if (c0 >= nrhs) return;
    const int nc = nrhs - c0 < CG ? nrhs - c0 : CG;
    const int64_t ld = F.p + F.u;
    const cpx<R>* L = factor + F.Loff;
    cpx<R>* yb = y + F.pstart + (int64_t)c0 * n;
    const int tid = threadIdx.x, lane = tid & 63, cb = (tid >> 6) * CPW;
    for (int64_t t = K0 + ((K1 - 1 - K0) / NB) * NB; t >= K0; t -= NB) {
        const int w = (int)(K1 - t < NB ? K1 - t : NB);
        for (int idx = tid; idx < NB * CG; idx += blockDim.x) {
            const int i = idx % NB, c = idx / NB;
            yt[i][c] = (c < nc && i < w) ? yb[(int64_t)c * n + t + i] : cpx<R>(R(0), R(0));
        }
        __syncthreads();
        if (lane < w) {
            const cpx<R>* Ui = tiles + F.tileoff + (t / NB) * 2 * NB * NB + NB * NB + lane;
            cpx<R> acc[CPW];
#pragma unroll
            for (int c = 0; c < CPW; ++c) acc[c] = cpx<R>(R(0), R(0));
            sp_dot<R, CG, CPW>(Ui, NB, w, yt, cb, acc);
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                xt[lane][cb + c] = acc[c];
                if (cb + c < nc) yb[(int64_t)(cb + c) * n + t + lane] = acc[c];
            }
        } else {
#pragma unroll
            for (int c = 0; c < CPW; ++c) xt[lane][cb + c] = cpx<R>(R(0), R(0));
        }
        __syncthreads();
        const int rows = (int)(t - K0);
        for (int r = lane; r < rows; r += 64) {
            const int64_t row = K0 + r;
            cpx<R> acc[CPW];
#pragma unroll
            for (int c = 0; c < CPW; ++c) acc[c] = cpx<R>(R(0), R(0));
            sp_dot<R, CG, CPW>(L + t * ld + row, ld, w, xt, cb, acc);
#pragma unroll
            for (int c = 0; c < CPW; ++c)
                if (cb + c < nc) {
                    cpx<R>* q = yb + (int64_t)(cb + c) * n + row;
                    *q = *q - acc[c];
                }
        }
        __syncthreads();
        if (t == K0) break;
    }
}

// The same super-panel solves for launches with few fronts (the top of the
// tree), where the sequential tile walk is the whole cost.  One 512-thread
// block per (front entry, group of CG columns): wave q owns the q-th 64-row
// tile of the super-panel (lane = row) and keeps its rows of the CG columns
// in registers.  At step t the wave owning tile t applies the tile inverse
// and publishes x_t through LDS; after one barrier the waves of the other
// tiles subtract their block L(q, t) x_t (U(q, t) going up), all 64 factor
// entries of a lane loaded at once.  The inverse of the next tile is staged
// in LDS by the whole block one step ahead, loaded into registers two steps
// ahead; the forward solve stores it with its columns in the pivot order
// (column pr[k] holds column k), so that the owner applies L_t^{-1} P_t to
// its rows as they stand.  A step's critical path is the owner's 64 x CG
// products, one barrier, one factor load and the next owner's 64 x CG
// products.  (Holding the next step's factor block in registers across the
// barrier needs 128 more registers per lane than 512-thread blocks leave.)
template <int CG>
__device__ __forceinline__ cpx<float> lane_value(const cpx<float> (&v)[CG], int c, int k)
{
    return cpx<float>(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[c].re), k)),
                      __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[c].im), k)));
}

template <int CG, bool FWD>
__global__ __launch_bounds__(512) void sp_wave_kernel(const wt_dfront* __restrict__ fr,
                                                      const int32_t* __restrict__ ent,
                                                      const cpx<float>* __restrict__ factor,
                                                      const cpx<float>* __restrict__ tiles,
                                                      const int32_t* __restrict__ perm, cpx<float>* __restrict__ y,
                                                      int64_t n, int nrhs)
{
    typedef cpx<float> C;
    constexpr int NB = WT_NB, NW = WT_NB2 / WT_NB, NTH = 64 * NW;
    constexpr int IPT = NB * NB / NTH;  // inverse entries staged per thread
    static_assert(NW == 8 && IPT * NTH == NB * NB, "one wave per tile of a super-panel");
    __shared__ C Is[2][NB * NB];  // [column][row]
    __shared__ C Xs[2][NB][CG];
    const int32_t* e = ent + 3 * blockIdx.y;
    const wt_dfront F = fr[e[0]];
    const int64_t K0 = e[1], K1 = e[2];
    const int c0 = blockIdx.x * CG;
    const int nc = nrhs - c0 < CG ? nrhs - c0 : CG;
    const int64_t ld = F.p + F.u;
    const C* L = factor + F.Loff;
    C* yb = y + F.pstart + (int64_t)c0 * n;
    const int tid = threadIdx.x, lane = tid & 63, q = tid >> 6;
    const int nt = (int)((K1 - K0 + NB - 1) / NB);
    const int64_t r0 = K0 + (int64_t)q * NB;
    const int wq = q < nt ? (int)(K1 - r0 < NB ? K1 - r0 : NB) : 0;
    const bool act = lane < wq;
    const int64_t row = r0 + lane;
    auto width = [&](int s) { return (int)(K1 - K0 - (int64_t)s * NB < NB ? K1 - K0 - (int64_t)s * NB : NB); };

    C iv[IPT];
    int id[IPT];  // LDS slot of iv
    auto load_inv = [&](int s) {
        const int64_t cb = K0 + (int64_t)s * NB;
        const int ws = width(s);
        const C* src = tiles + F.tileoff + (cb / NB) * 2 * NB * NB + (FWD ? 0 : NB * NB);
        // every load in bounds (clamped), the entries past the tile width zeroed
#pragma unroll
        for (int m = 0; m < IPT; ++m) {
            const int j = tid + m * NTH, k = j / NB, i = j % NB;
            const bool in = k < ws && i < ws;
            const C v = src[j];
            iv[m] = in ? v : C(0.f, 0.f);
            if (FWD) {
                const int pk = perm[F.pivoff + cb + (k < ws ? k : ws - 1)];
                id[m] = (in ? pk : k) * NB + i;
            } else {
                id[m] = j;
            }
        }
    };
    auto store_inv = [&](int buf) {
#pragma unroll
        for (int m = 0; m < IPT; ++m) Is[buf][id[m]] = iv[m];
    };
    C a[NB];  // this wave's factor block for the next update
    auto fetch = [&](int s) {
        const int64_t cb = K0 + (int64_t)s * NB;
        const int ws = width(s);
        const C* src = L + cb * ld + (act ? row : r0);
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const C v = src[(int64_t)(k < ws ? k : ws - 1) * ld];
            a[k] = (act && k < ws) ? v : C(0.f, 0.f);
        }
    };
    C acc[CG];
#pragma unroll
    for (int c = 0; c < CG; ++c) acc[c] = (act && c < nc) ? yb[(int64_t)c * n + row] : C(0.f, 0.f);

    auto tile_at = [&](int st) { return FWD ? st : nt - 1 - st; };
    load_inv(tile_at(0));
    store_inv(0);
    if (nt > 1) load_inv(tile_at(1));
    __syncthreads();
    for (int st = 0; st < nt; ++st) {
        const int t = tile_at(st), buf = st & 1;
        if (st + 1 < nt) {
            store_inv(buf ^ 1);
            if (st + 2 < nt) load_inv(tile_at(st + 2));
        }
        if (q == t) {
            C x[CG];
#pragma unroll
            for (int c = 0; c < CG; ++c) x[c] = C(0.f, 0.f);
            const C* I = Is[buf];
#pragma unroll 4
            for (int k = 0; k < NB; ++k) {
                const C lk = I[k * NB + lane];
#pragma unroll
                for (int c = 0; c < CG; ++c) cmac(x[c], lk, lane_value<CG>(acc, c, k));
            }
#pragma unroll
            for (int c = 0; c < CG; ++c) {
                Xs[buf][lane][c] = x[c];
                if (act && c < nc) yb[(int64_t)c * n + row] = x[c];
            }
        }
        __syncthreads();
        if (FWD ? (q > t && q < nt) : (q < t)) {
            fetch(t);
#pragma unroll
            for (int k = 0; k < NB; ++k) {
#pragma unroll
                for (int c = 0; c < CG; ++c) cmsub(acc[c], a[k], Xs[buf][k][c]);
            }
        }
    }
}

// ---- solve-side gathers / scatters (block of nrhs columns) ----

// y (elimination order, ld = n) += children's update rows that land on the
// parent's pivots; the parent's own update block W gets the rest.
template <typename R>
__global__ __launch_bounds__(256) void solve_gather_kernel(const wt_dfront* __restrict__ fr,
                                                           const int32_t* __restrict__ kids,
                                                           const int32_t* __restrict__ relmap,
                                                           const cpx<R>* __restrict__ w_child,
                                                           cpx<R>* __restrict__ w_parent, cpx<R>* __restrict__ y,
                                                           int64_t n, int nrhs)
{
    const wt_dfront C = fr[kids[blockIdx.x]];
    const wt_dfront P = fr[C.parent];
    const int32_t* map = relmap + C.moff;
    const int64_t uc = C.u;
    const int r = blockIdx.y;
    if (r >= nrhs) return;
    const cpx<R>* wc = w_child + C.Woff * nrhs + (int64_t)r * uc;
    cpx<R>* wp = w_parent + P.Woff * nrhs + (int64_t)r * P.u;
    for (int64_t i = threadIdx.x; i < uc; i += blockDim.x) {
        int64_t pos = map[i];
        cpx<R> v = wc[i];
        if (pos < P.p) {
            cpx<R>* t = y + (int64_t)r * n + P.pstart + pos;
            *t = *t + v;
        } else {
            cpx<R>* t = wp + (pos - P.p);
            *t = *t + v;
        }
    }
}

// X_U (u x nrhs work block) = x[upd_idx of the front] for the back solve.
template <typename R>
__global__ __launch_bounds__(256) void solve_gather_upd_kernel(const wt_dfront* __restrict__ fr,
                                                               const int32_t* __restrict__ list,
                                                               const int64_t* __restrict__ upd_idx,
                                                               const cpx<R>* __restrict__ x,
                                                               cpx<R>* __restrict__ work, int64_t n, int nrhs)
{
    const wt_dfront F = fr[list[blockIdx.x]];
    const int r = blockIdx.y;
    if (r >= nrhs) return;
    const int64_t* idx = upd_idx + F.uoff;
    cpx<R>* wk = work + F.Woff * nrhs + (int64_t)r * F.u;
    for (int64_t i = threadIdx.x; i < F.u; i += blockDim.x) wk[i] = x[(int64_t)r * n + idx[i]];
}

}  // namespace

static __global__ void zero_i32_kernel(int* __restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

static __global__ void pivstat_init_kernel(int* __restrict__ p)
{
    const int i = threadIdx.x;
    if (i < WT_PIVSTAT_LEN) p[i] = i == 2 ? 0x7f800000 : 0;  // [2] starts at +inf
}

int wt_pivstat_init(int* p, hipStream_t s)
{
    hipLaunchKernelGGL(pivstat_init_kernel, dim3(1), dim3(64), 0, s, p);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

// A fill kernel rather than a memset node, so that captured sequences
// replay it as an ordinary kernel.
int wt_zero_i32(int* p, int n, hipStream_t s)
{
    if (n <= 0) return WT_OK;
    hipLaunchKernelGGL(zero_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template <typename R>
int wt_fill_zero(cpx<R>* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return WT_OK;
    const int64_t bytes = n * (int64_t)sizeof(cpx<R>);
    if (FILL16 && ((uintptr_t)p & 15) == 0 && bytes >= 16) {
        // the 16-byte units, then an odd complex64 entry at the end
        const int64_t n16 = bytes / 16;
        int64_t blocks = (n16 + 1023) / 1024;
        if (blocks > 16384) blocks = 16384;
        hipLaunchKernelGGL(fill_zero16_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (uint4*)p, n16);
        WT_HIP(hipGetLastError());
        const int64_t done = n16 * 16 / (int64_t)sizeof(cpx<R>);
        if (done == n) return WT_OK;
        p += done;
        n -= done;
    }
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_zero_kernel<R>, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template <typename Ri, typename R>
int wt_scatter(const cpx<Ri>* vals, const int64_t* dst, int64_t nnz, cpx<R>* arena, hipStream_t s)
{
    if (nnz <= 0) return WT_OK;
    int64_t blocks = (nnz + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((scatter_kernel<Ri, R>), dim3((unsigned)blocks), dim3(256), 0, s, vals, dst, nnz, arena);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

int wt_row_scale(const int32_t* rowptr, const cpx<double>* vals, int64_t n, double* scale, hipStream_t s)
{
    if (n <= 0) return WT_OK;
    hipLaunchKernelGGL(row_scale_kernel<double>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, rowptr, vals, n,
                       scale);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template <typename R>
int wt_scatter_rows(const int32_t* rowptr, const cpx<double>* vals, const int64_t* dst, const double* scale,
                    int64_t n, cpx<R>* arena, hipStream_t s)
{
    if (n <= 0) return WT_OK;
    hipLaunchKernelGGL((scatter_rows_kernel<double, R>), dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, rowptr,
                       vals, dst, scale, n, arena);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template <typename R>
int wt_extend_add(const wt_dfront* fr, const int32_t* kids, const int32_t* col_start, int nkids, int64_t nblocks,
                  const int32_t* relmap, const cpx<R>* cb_child, cpx<R>* factor, cpx<R>* cb_parent, int part,
                  hipStream_t s)
{
    if (nkids <= 0 || nblocks <= 0) return WT_OK;
    if (part < 0 || part > 2 || nblocks > 0x7fffffff) return WT_ERR_ARG;
    hipLaunchKernelGGL(extend_add_kernel<R>, dim3((unsigned)nblocks), dim3(256), 0, s, fr, kids, col_start, nkids,
                       relmap, cb_child, factor, cb_parent, part);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template <typename R>
int wt_diag_tiles(const wt_dfront* fr, const int32_t* list, int nlist, int64_t kb, cpx<R>* factor, cpx<R>* tiles,
                  int32_t* perm, int* status, hipStream_t s)
{
    if (nlist <= 0) return WT_OK;
    if constexpr (sizeof(R) == 4)
        hipLaunchKernelGGL(diag_tile_wave_kernel, dim3(nlist), dim3(128), 0, s, fr, list, kb, factor, tiles, perm,
                           status);
    else
        hipLaunchKernelGGL(diag_tile_kernel<R>, dim3(nlist), dim3(256), 0, s, fr, list, kb, factor, tiles, perm,
                           status);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template <typename R>
int wt_solve_gather(const wt_dfront* fr, const int32_t* kids, int nkids, const int32_t* relmap,
                    const cpx<R>* w_child, cpx<R>* w_parent, cpx<R>* y, int64_t n, int nrhs, hipStream_t s)
{
    if (nkids <= 0) return WT_OK;
    hipLaunchKernelGGL(solve_gather_kernel<R>, dim3(nkids, nrhs), dim3(256), 0, s, fr, kids, relmap, w_child,
                       w_parent, y, n, nrhs);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template <typename R>
int wt_solve_gather_upd(const wt_dfront* fr, const int32_t* list, int nlist, const int64_t* upd_idx,
                        const cpx<R>* x, cpx<R>* work, int64_t n, int nrhs, hipStream_t s)
{
    if (nlist <= 0) return WT_OK;
    hipLaunchKernelGGL(solve_gather_upd_kernel<R>, dim3(nlist, nrhs), dim3(256), 0, s, fr, list, upd_idx, x, work,
                       n, nrhs);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template int wt_scatter_rows<float>(const int32_t*, const cpx<double>*, const int64_t*, const double*, int64_t,
                                    cpx<float>*, hipStream_t);
template int wt_scatter_rows<double>(const int32_t*, const cpx<double>*, const int64_t*, const double*, int64_t,
                                     cpx<double>*, hipStream_t);

template <typename R>
int wt_sp_solve(int fwd, const wt_dfront* fr, const int32_t* ent, int nent, const cpx<R>* factor,
                const cpx<R>* tiles, const int32_t* perm, cpx<R>* y, int64_t n, int nrhs, hipStream_t s)
{
    if (nent <= 0) return WT_OK;
    // 16-column groups load each factor entry for the most columns; launches
    // with few fronts (the top of the tree) use 4-column groups so that more
    // blocks share the sequential tile walk
    const bool wide = (int64_t)nent * ((nrhs + 15) / 16) >= 1024;
    if constexpr (sizeof(R) == 4)
    if (!wide) {
        // one wave per tile; fewer columns per block where fewer fronts share the launch
        const int cg = nent <= 2 ? 1 : nent <= 8 ? 2 : 4;
        const dim3 g2((nrhs + cg - 1) / cg, nent);
        if (nent > 65535) return WT_ERR_ARG;
        const cpx<float>* fa = (const cpx<float>*)factor;
        const cpx<float>* ta = (const cpx<float>*)tiles;
        cpx<float>* ya = (cpx<float>*)y;
#define WT_SPW(CG, FW) \
    hipLaunchKernelGGL((sp_wave_kernel<CG, FW>), g2, dim3(512), 0, s, fr, ent, fa, ta, perm, ya, n, nrhs)
        if (fwd) {
            if (cg == 1) WT_SPW(1, true);
            else if (cg == 2) WT_SPW(2, true);
            else WT_SPW(4, true);
        } else {
            if (cg == 1) WT_SPW(1, false);
            else if (cg == 2) WT_SPW(2, false);
            else WT_SPW(4, false);
        }
#undef WT_SPW
        WT_HIP(hipGetLastError());
        return WT_OK;
    }
    const int cg = wide ? 16 : 4;
    dim3 grid(nent, (nrhs + cg - 1) / cg);
    if (fwd) {
        if (wide)
            hipLaunchKernelGGL((sp_fwd_kernel<R, 16>), grid, dim3(256), 0, s, fr, ent, factor, tiles, perm, y, n, nrhs);
        else
            hipLaunchKernelGGL((sp_fwd_kernel<R, 4>), grid, dim3(256), 0, s, fr, ent, factor, tiles, perm, y, n, nrhs);
    } else {
        if (wide)
            hipLaunchKernelGGL((sp_bwd_kernel<R, 16>), grid, dim3(256), 0, s, fr, ent, factor, tiles, y, n, nrhs);
        else
            hipLaunchKernelGGL((sp_bwd_kernel<R, 4>), grid, dim3(256), 0, s, fr, ent, factor, tiles, y, n, nrhs);
    }
    WT_HIP(hipGetLastError());
    return WT_OK;
}
template int wt_sp_solve<float>(int, const wt_dfront*, const int32_t*, int, const cpx<float>*, const cpx<float>*,
                                const int32_t*, cpx<float>*, int64_t, int, hipStream_t);
template int wt_sp_solve<double>(int, const wt_dfront*, const int32_t*, int, const cpx<double>*,
                                 const cpx<double>*, const int32_t*, cpx<double>*, int64_t, int, hipStream_t);

#define WT_INST(R)                                                                                              \
    template int wt_fill_zero<R>(cpx<R>*, int64_t, hipStream_t);                                                \
    template int wt_extend_add<R>(const wt_dfront*, const int32_t*, const int32_t*, int, int64_t, const int32_t*, \
                                  const cpx<R>*, cpx<R>*, cpx<R>*, int, hipStream_t);                           \
    template int wt_diag_tiles<R>(const wt_dfront*, const int32_t*, int, int64_t, cpx<R>*, cpx<R>*, int32_t*,   \
                                  int*, hipStream_t);                                                           \
    template int wt_solve_gather<R>(const wt_dfront*, const int32_t*, int, const int32_t*, const cpx<R>*,       \
                                    cpx<R>*, cpx<R>*, int64_t, int, hipStream_t);                               \
    template int wt_solve_gather_upd<R>(const wt_dfront*, const int32_t*, int, const int64_t*, const cpx<R>*,   \
                                        cpx<R>*, int64_t, int, hipStream_t);
WT_INST(float)
WT_INST(double)
template int wt_scatter<double, float>(const cpx<double>*, const int64_t*, int64_t, cpx<float>*, hipStream_t);
template int wt_scatter<double, double>(const cpx<double>*, const int64_t*, int64_t, cpx<double>*, hipStream_t);
template int wt_scatter<float, float>(const cpx<float>*, const int64_t*, int64_t, cpx<float>*, hipStream_t);
