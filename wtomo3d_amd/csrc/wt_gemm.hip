// Batched complex GEMM for the dense front kernels, on CDNA4 matrix cores.
//
// C (op)= A * B for many independent problems in one launch (a flattened
// list of output tiles).  Complex products use the 3M form: three real
// MFMA products T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi) with fp32
// (or fp64) accumulation, then Re = T1 - T2, Im = T3 - T1 - T2.
//
// The MFMA computes D = X Y with X taken from B^T and Y from A^T, so D is a
// tile of C^T: each lane's accumulators then run along the rows of the
// column-major C, and the epilogue stores are contiguous across lanes.
// 256 threads = 4 waves; fp32 uses v_mfma_f32_32x32x2f32 (default block
// 128 x 128 as 2 x 2 waves of 64 x 64), fp64 uses v_mfma_f64_16x16x4f64
// (the solver's default for fp64 is the 64 x 64 block, two per CU); the
// tile variants are
// listed with wt_gemm_tile.  K is consumed in slices of 16 complex values,
// staged through double-buffered LDS as separate real and imaginary planes.
#include "wt_gemm.h"

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename R>
struct MF;
template <>
struct MF<float> {
    static constexpr int MT = 32, KS = 2;
    typedef f32x16 acc_t;
    static constexpr int NACC = 16;
    __device__ static acc_t mfma(float x, float y, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(x, y, c, 0, 0, 0);
    }
    // accumulator v of lane l holds D[drow][dcol]
    __device__ static int drow(int l, int v) { return (v >> 2) * 8 + (l >> 5) * 4 + (v & 3); }
    __device__ static int dcol(int l) { return l & 31; }
    // operand element of lane l: row/col index and k index inside the step
    __device__ static int orow(int l) { return l & 31; }
    __device__ static int ok(int l) { return l >> 5; }
};
template <>
struct MF<double> {
    static constexpr int MT = 16, KS = 4;
    typedef f64x4 acc_t;
    static constexpr int NACC = 4;
    __device__ static acc_t mfma(double x, double y, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
    }
    __device__ static int drow(int l, int v) { return v * 4 + (l >> 4); }
    __device__ static int dcol(int l) { return l & 15; }
    __device__ static int orow(int l) { return l & 15; }
    __device__ static int ok(int l) { return l >> 4; }
};

// The problem list lives in device memory, so its operand pointers reach the
// kernel as generic pointers and would be accessed with flat instructions,
// which the hardware counts with the LDS operations: every wait for an LDS
// read would then also wait for the prefetch of the next K slice.  Operands
// are therefore accessed through global-address-space pointers to 2-wide
// vectors (one complex value each), which compile to global loads and
// stores whose completion is tracked separately.
template <typename R>
using v2 = R __attribute__((ext_vector_type(2)));
template <typename R>
using gv2 = __attribute__((address_space(1))) v2<R>;
typedef __attribute__((address_space(1))) const int32_t gi32;

constexpr int NT = 256;  // threads per block
#ifndef GEMM_CPREF
#define GEMM_CPREF 1
#endif
// complex k per slice.  The slice depth does not change the order in which
// any accumulator sums its k terms (the MFMA sequence over k is the same),
// so results do not depend on it; 16 halves the barriers per k of 8 and
// measured 14-26 % faster for fp64 and 2-9 % for fp32.
constexpr int BK = 16;

// Blocks per CU the register allocation is sized for.  fp32: 2.  fp64: 2
// for the 64 x 64 tile, whose 70 KB of LDS lets two blocks share a CU once
// each wave fits in 256 registers (measured 1.5x the single-block rate:
// 75 against 48 TF/s at 4096^3); 1 for the 128-row tiles, whose 100 KB of
// LDS allows only one.
template <typename R, int V>
constexpr int gemm_occupancy()
{
    return sizeof(R) == 4 ? 2 : (V == 2 ? 2 : 1);
}

template <typename R, int V, bool CP = false>
__global__ __launch_bounds__(NT, (gemm_occupancy<R, V>())) void gemm_batched_kernel(const wt_gemm_prob<R>* __restrict__ probs,
                                                            const int64_t* __restrict__ tile_start,
                                                            int nprob, int64_t ntiles)
{
    typedef MF<R> M_;
    typedef wt_gemm_tile<R, V> TL;
    constexpr int BM = TL::BM, BN = TL::BN;
    constexpr int WM = BM / TL::WGM, WN = BN / TL::WGN;
    constexpr int SM = WM / M_::MT, SN = WN / M_::MT;
    constexpr int EA = BM * BK / NT, EB = BK * BN / NT;
    constexpr int TPA = BM / EA, TPB = BK / EB;  // threads per column of A / B
    constexpr int PAD = 4;

    __shared__ R sAr[2][BK][BM + PAD], sAi[2][BK][BM + PAD];
    __shared__ R sBr[2][BK][BN + PAD], sBi[2][BK][BN + PAD];

    // tile id with an XCD-contiguous remap (blocks go round-robin to 8 XCDs)
    int64_t b = blockIdx.x;
    {
        int64_t q = ntiles / 8, r = ntiles % 8, x = b % 8, i = b / 8;
        b = x * q + (x < r ? x : r) + i;
    }
    // problem of this tile
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const wt_gemm_prob<R> P = probs[lo];
    const gv2<R>* gA = (const gv2<R>*)P.A;
    const gv2<R>* gB = (const gv2<R>*)P.B;
    gv2<R>* gC = (gv2<R>*)P.C;
    gi32* gperm = (gi32*)P.bperm;
    int64_t t = b - tile_start[lo];
    const int tiles_m = (P.M + BM - 1) / BM, tiles_n = (P.N + BN - 1) / BN;
    // grouped order: GM tile-rows at a time share B columns in cache
    constexpr int GM = 8;
    int64_t grp = t / ((int64_t)GM * tiles_n);
    int first_m = (int)(grp * GM);
    int gm = tiles_m - first_m < GM ? tiles_m - first_m : GM;
    int64_t tin = t - grp * (int64_t)GM * tiles_n;
    const int tm = first_m + (int)(tin % gm), tn = (int)(tin / gm);
    const int m0 = tm * BM, n0 = tn * BN;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % TL::WGM) * WM, wn0 = (wave / TL::WGM) * WN;

    typename M_::acc_t acc[3][SM][SN];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int i = 0; i < SM; ++i)
#pragma unroll
            for (int j = 0; j < SN; ++j)
#pragma unroll
                for (int v = 0; v < M_::NACC; ++v) acc[c][i][j][v] = R(0);

    const int K = P.K;
    const int nkt = (K + BK - 1) / BK;
    // per-thread load coordinates
    const int ka = tid / TPA, ma = (tid % TPA) * EA;
    const int nbb = tid / TPB, kb = (tid % TPB) * EB;
    v2<R> ra[EA], rb[EB];

    auto load = [&](int kt) {
        const int k0 = kt * BK;
        const bool fullA = (m0 + BM <= P.M) && (k0 + BK <= K);
        const int kk = k0 + ka;
        const gv2<R>* pa = gA + (int64_t)kk * P.lda + m0 + ma;
        if (fullA) {
#pragma unroll
            for (int e = 0; e < EA; ++e) ra[e] = pa[e];
        } else {
#pragma unroll
            for (int e = 0; e < EA; ++e)
                ra[e] = (kk < K && m0 + ma + e < P.M) ? pa[e] : v2<R>{R(0), R(0)};
        }
        const int nn = n0 + nbb;
        const bool fullB = (n0 + BN <= P.N) && (k0 + BK <= K) && !gperm;
        if (fullB) {
            const gv2<R>* pb = gB + (int64_t)nn * P.ldb + k0 + kb;
#pragma unroll
            for (int e = 0; e < EB; ++e) rb[e] = pb[e];
        } else {
#pragma unroll
            for (int e = 0; e < EB; ++e) {
                int k = k0 + kb + e;
                if (k < K && nn < P.N) {
                    int kr = gperm ? gperm[k] : k;
                    rb[e] = gB[(int64_t)nn * P.ldb + kr];
                } else {
                    rb[e] = v2<R>{R(0), R(0)};
                }
            }
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < EA; ++e) { sAr[buf][ka][ma + e] = ra[e].x; sAi[buf][ka][ma + e] = ra[e].y; }
#pragma unroll
        for (int e = 0; e < EB; ++e) { sBr[buf][kb + e][nbb] = rb[e].x; sBi[buf][kb + e][nbb] = rb[e].y; }
    };

    // CP (launches whose products all read C: C -= A B, C += A B): C loaded
    // with the first slice, so its latency overlaps the K loop -- an instance
    // of its own, as the 32 registers it holds per lane would cost the other
    // launches occupancy (tiles whose lanes hold at most 16 outputs)
    const int mode = P.mode;
    constexpr bool CPRE = CP && GEMM_CPREF && SM * SN * M_::NACC <= 16;
    const bool rmw = mode == WT_GEMM_SUB || mode == WT_GEMM_ADD;
    v2<R> cpre[CPRE ? SM : 1][CPRE ? SN : 1][CPRE ? M_::NACC : 1];
    if constexpr (CPRE) {
        if (rmw) {
#pragma unroll
            for (int i = 0; i < SM; ++i)
#pragma unroll
                for (int j = 0; j < SN; ++j)
#pragma unroll
                    for (int v = 0; v < M_::NACC; ++v) {
                        const int m = m0 + wm0 + i * M_::MT + M_::dcol(lane);
                        const int n = n0 + wn0 + j * M_::MT + M_::drow(lane, v);
                        cpre[i][j][v] = (m < P.M && n < P.N) ? gC[(int64_t)n * P.ldc + m] : v2<R>{R(0), R(0)};
                    }
        }
    }
    if (nkt > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    const int orow = M_::orow(lane), okk = M_::ok(lane);
    for (int kt = 0; kt < nkt; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nkt) load(kt + 1);
#pragma unroll
        for (int ks = 0; ks < BK; ks += M_::KS) {
            R xr[SN], xi[SN], xs[SN], yr[SM], yi[SM], ys[SM];
#pragma unroll
            for (int j = 0; j < SN; ++j) {
                xr[j] = sBr[cur][ks + okk][wn0 + j * M_::MT + orow];
                xi[j] = sBi[cur][ks + okk][wn0 + j * M_::MT + orow];
                xs[j] = xr[j] + xi[j];
            }
#pragma unroll
            for (int i = 0; i < SM; ++i) {
                yr[i] = sAr[cur][ks + okk][wm0 + i * M_::MT + orow];
                yi[i] = sAi[cur][ks + okk][wm0 + i * M_::MT + orow];
                ys[i] = yr[i] + yi[i];
            }
#pragma unroll
            for (int i = 0; i < SM; ++i)
#pragma unroll
                for (int j = 0; j < SN; ++j) {
                    acc[0][i][j] = M_::mfma(xr[j], yr[i], acc[0][i][j]);
                    acc[1][i][j] = M_::mfma(xi[j], yi[i], acc[1][i][j]);
                    acc[2][i][j] = M_::mfma(xs[j], ys[i], acc[2][i][j]);
                }
        }
        if (kt + 1 < nkt) store(cur ^ 1);
        __syncthreads();
    }

    // epilogue
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
        for (int j = 0; j < SN; ++j)
#pragma unroll
            for (int v = 0; v < M_::NACC; ++v) {
                const int m = m0 + wm0 + i * M_::MT + M_::dcol(lane);
                const int n = n0 + wn0 + j * M_::MT + M_::drow(lane, v);
                if (m < P.M && n < P.N) {
                    R t1 = acc[0][i][j][v], t2 = acc[1][i][j][v], t3 = acc[2][i][j][v];
                    R re = t1 - t2, im = (t3 - t1) - t2;
                    gv2<R>* pc = gC + (int64_t)n * P.ldc + m;
                    if (mode == WT_GEMM_SET) {
                        *pc = v2<R>{re, im};
                    } else if (mode == WT_GEMM_NSET) {
                        *pc = v2<R>{-re, -im};
                    } else {
                        v2<R> c;
                        if constexpr (CPRE) c = cpre[i][j][v];
                        else c = *pc;
                        if (mode == WT_GEMM_SUB) *pc = v2<R>{c.x - re, c.y - im};
                        else *pc = v2<R>{c.x + re, c.y + im};
                    }
                }
            }
}

// ---- fp32 products on the bf16 matrix cores (tile variant 4) ----
//
// Each fp32 operand x is split into three bf16 parts, x = h + m + l with
// h = bf16(x), m = bf16(x - h), l = bf16(x - h - m), each rounded to
// nearest even (both differences are exact in fp32), so the parts carry the
// 24 bits of x's significand.  A real product a b is formed from the six
// pairs of the split parts down to order 2^-16 |a b|,
//   a b = h h' + (h m' + m h') + (h l' + m m' + l h') + O(2^-23 |a b|),
// each pair exact in fp32 and accumulated in fp32 by v_mfma_f32_32x32x16_bf16,
// so a product carries about the error of one fp32 rounding.  Six bf16
// MFMAs cost 6 x 32768 / 2.5 PFLOP/s against 8 x 4096 / 157 TFLOP/s for the
// same 32 x 32 x 16 real product on the fp32 matrix cores (2.7x less time at
// equal utilisation).  The 3M form is kept: the planes are Re, Im and
// Re + Im (the sum rounded in fp32, then split), nine bf16 planes per
// operand.
//
// Geometry: 256 threads, a 128 x 128 tile as 2 x 2 waves of 64 x 64 (2 x 2
// MFMA blocks of 32 x 32 each, three accumulators per block: T1 = Re Re,
// T2 = Im Im, T3 = Sum Sum), K in slices of 16 = one MFMA depth.  A slice
// is staged in LDS as [plane][row][16 k] with k contiguous, which is the
// MFMA operand order (lane l reads 8 consecutive k of row l % 32), double
// buffered: 2 x 2 x 9 x 128 x 16 bf16 = 144 KB, one block per CU.  The two
// 16-byte halves of a row swap places every 4 rows, so that the 16-byte
// reads and writes of 8 consecutive rows cover all 32 banks.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// The split of two values at once, each part of both packed in one dword
// (element 0 in the low half): one pair conversion per part, and the parts
// come back to fp32 by a shift (low) or a mask (high) -- 11 VALU operations
// per pair against 7 per value split one at a time (4096^3: 193 against
// 181 TF/s fp32-equivalent, identical parts).
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t){a, b}, bf16x2_t));
}

__device__ __forceinline__ void split3x2(float x0, float x1, uint32_t& h, uint32_t& m, uint32_t& l)
{
    h = pack_bf16x2(x0, x1);
    const float r0 = x0 - __uint_as_float(h << 16), r1 = x1 - __uint_as_float(h & 0xffff0000u);
    m = pack_bf16x2(r0, r1);
    l = pack_bf16x2(r0 - __uint_as_float(m << 16), r1 - __uint_as_float(m & 0xffff0000u));
}

// 16-byte chunk (0 or 1) of a row's 16 k as stored: swapped every 4 rows
__device__ __forceinline__ int split_chunk(int row, int chunk) { return chunk ^ ((row >> 2) & 1); }

// T x T output tile (T = 128: waves of 64 x 64, one block per CU, every
// thread stages one A row and one B column of a slice; T = 64: waves of
// 32 x 32, two blocks per CU, threads 0-127 stage A and 128-255 stage B).
template <int T>
__global__ __launch_bounds__(NT, T == 128 ? 1 : 2) void gemm_split_kernel(const wt_gemm_prob<float>* __restrict__ probs,
                                                                          const int64_t* __restrict__ tile_start,
                                                                          int nprob, int64_t ntiles)
{
    typedef MF<float> M_;
    constexpr int BM = T, BN = T, KC = 16, NP = 9;
    constexpr int SB = T / 64;           // 32 x 32 MFMA blocks per wave and dimension
    constexpr bool BOTH = T == 128;      // every thread stages both operands
    __shared__ __align__(16) __bf16 sA[2][NP][BM][KC];
    __shared__ __align__(16) __bf16 sB[2][NP][BN][KC];

    int64_t b = blockIdx.x;
    {
        int64_t q = ntiles / 8, r = ntiles % 8, x = b % 8, i = b / 8;
        b = x * q + (x < r ? x : r) + i;
    }
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const wt_gemm_prob<float> P = probs[lo];
    const gv2<float>* gA = (const gv2<float>*)P.A;
    const gv2<float>* gB = (const gv2<float>*)P.B;
    gv2<float>* gC = (gv2<float>*)P.C;
    gi32* gperm = (gi32*)P.bperm;
    int64_t t = b - tile_start[lo];
    const int tiles_m = (P.M + BM - 1) / BM, tiles_n = (P.N + BN - 1) / BN;
    constexpr int GM = 8;
    int64_t grp = t / ((int64_t)GM * tiles_n);
    int first_m = (int)(grp * GM);
    int gm = tiles_m - first_m < GM ? tiles_m - first_m : GM;
    int64_t tin = t - grp * (int64_t)GM * tiles_n;
    const int tm = first_m + (int)(tin % gm), tn = (int)(tin / gm);
    const int m0 = tm * BM, n0 = tn * BN;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave & 1) * (T / 2), wn0 = (wave >> 1) * (T / 2);

    f32x16 acc[3][SB][SB];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int i = 0; i < SB; ++i)
#pragma unroll
            for (int j = 0; j < SB; ++j)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[c][i][j][v] = 0.f;

    const int K = P.K, nkt = (K + KC - 1) / KC;
    // loader: row (of A) / column (of B) lr of the tile, k chunk kh (8 k);
    // with !BOTH, isb selects the operand this thread stages
    const int lr = tid & (T - 1), kh = BOTH ? tid >> 7 : (tid >> 6) & 1;
    const bool isb = !BOTH && tid >= 128;
    v2<float> ra[8], rb[8];
    auto load_a = [&](int k0, v2<float> (&r)[8]) {
        const int m = m0 + lr;
        const gv2<float>* pa = gA + (int64_t)k0 * P.lda + m;
        if (m < P.M && k0 + 8 <= K) {
#pragma unroll
            for (int e = 0; e < 8; ++e) r[e] = pa[(int64_t)e * P.lda];
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e)
                r[e] = (m < P.M && k0 + e < K) ? pa[(int64_t)e * P.lda] : v2<float>{0.f, 0.f};
        }
    };
    auto load_b = [&](int k0, v2<float> (&r)[8]) {
        const int n = n0 + lr;
        if (n < P.N && k0 + 8 <= K && !gperm) {
            const gv2<float>* pb = gB + (int64_t)n * P.ldb + k0;
#pragma unroll
            for (int e = 0; e < 8; ++e) r[e] = pb[e];
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = k0 + e;
                r[e] = (k < K && n < P.N) ? gB[(int64_t)n * P.ldb + (gperm ? gperm[k] : k)] : v2<float>{0.f, 0.f};
            }
        }
    };
    auto load = [&](int kt) {
        const int k0 = kt * KC + kh * 8;
        if (BOTH) {
            load_a(k0, ra);
            load_b(k0, rb);
        } else if (isb) {
            load_b(k0, ra);
        } else {
            load_a(k0, ra);
        }
    };
    // one value kind (Re, Im, Re + Im) of one operand at a time, its three
    // planes written before the next kind is split
    auto store_kinds = [&](const v2<float> (&r)[8], __bf16 (*dst)[BM][KC]) {
        const int c = split_chunk(lr, kh) * 8;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            u32x4 ph, pm, pl;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const v2<float> a = r[2 * e], b = r[2 * e + 1];
                const float x0 = q == 0 ? a.x : (q == 1 ? a.y : a.x + a.y);
                const float x1 = q == 0 ? b.x : (q == 1 ? b.y : b.x + b.y);
                uint32_t h, m, l;
                split3x2(x0, x1, h, m, l);
                ph[e] = h;
                pm[e] = m;
                pl[e] = l;
            }
            *(u32x4*)&dst[3 * q][lr][c] = ph;
            *(u32x4*)&dst[3 * q + 1][lr][c] = pm;
            *(u32x4*)&dst[3 * q + 2][lr][c] = pl;
        }
    };
    auto store = [&](int buf) {
        if (BOTH) {
            store_kinds(ra, sA[buf]);
            store_kinds(rb, sB[buf]);
        } else {
            store_kinds(ra, isb ? sB[buf] : sA[buf]);
        }
    };

    if (nkt > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    const int orow = lane & 31, oc = split_chunk(orow, lane >> 5) * 8;
    // Per slice: the next slice's global loads, this slice's fragment reads
    // and MFMAs, then the next slice's split and LDS writes.  The split needs
    // the loads, which arrive while the MFMAs run; interleaving it among the
    // MFMAs (a loop with the last slice peeled, so that both are one basic
    // block) made the MFMAs wait for the loads and measured 3-8 % slower.
    for (int kt = 0; kt < nkt; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nkt) load(kt + 1);
        // X = B^T (rows n), Y = A^T (columns m): D = X Y is a tile of C^T
        bf16x8 x[SB][NP], y[SB][NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
#pragma unroll
            for (int j = 0; j < SB; ++j) x[j][p] = *(const bf16x8*)&sB[cur][p][wn0 + 32 * j + orow][oc];
#pragma unroll
            for (int i = 0; i < SB; ++i) y[i][p] = *(const bf16x8*)&sA[cur][p][wm0 + 32 * i + orow][oc];
        }
        // pairs (x part, y part), smallest first; 3 SB^2 independent
        // accumulators between two updates of one
        constexpr int PX[6] = {0, 1, 2, 0, 1, 0}, PY[6] = {2, 1, 0, 1, 0, 0};
#pragma unroll
        for (int s = 0; s < 6; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int i = 0; i < SB; ++i)
#pragma unroll
                    for (int j = 0; j < SB; ++j)
                        acc[c][i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x[j][3 * c + PX[s]],
                                                                               y[i][3 * c + PY[s]], acc[c][i][j], 0,
                                                                               0, 0);
        if (kt + 1 < nkt) store(cur ^ 1);
        __syncthreads();
    }

    const int mode = P.mode;
#pragma unroll
    for (int i = 0; i < SB; ++i)
#pragma unroll
        for (int j = 0; j < SB; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int m = m0 + wm0 + i * 32 + M_::dcol(lane);
                const int n = n0 + wn0 + j * 32 + M_::drow(lane, v);
                if (m < P.M && n < P.N) {
                    const float t1 = acc[0][i][j][v], t2 = acc[1][i][j][v], t3 = acc[2][i][j][v];
                    const float re = t1 - t2, im = (t3 - t1) - t2;
                    gv2<float>* pc = gC + (int64_t)n * P.ldc + m;
                    if (mode == WT_GEMM_SET) {
                        *pc = v2<float>{re, im};
                    } else if (mode == WT_GEMM_NSET) {
                        *pc = v2<float>{-re, -im};
                    } else {
                        const v2<float> c = *pc;
                        if (mode == WT_GEMM_SUB) *pc = v2<float>{c.x - re, c.y - im};
                        else *pc = v2<float>{c.x + re, c.y + im};
                    }
                }
            }
}

}  // namespace

template <typename R>
int wt_gemm_launch(const wt_gemm_prob<R>* d_probs, const int64_t* d_tile_start, int nprob, int64_t ntiles, int var,
                   hipStream_t stream, int rmw)
{
    if (nprob <= 0 || ntiles <= 0) return WT_OK;
    if (ntiles > 0x7fffffff) return WT_ERR_ARG;
    // the 64 x 64 tile with C loaded ahead, for launches that read C
    auto tile64 = [&]() {
        if (rmw)
            hipLaunchKernelGGL((gemm_batched_kernel<R, 2, true>), dim3((unsigned)ntiles), dim3(NT), 0, stream,
                               d_probs, d_tile_start, nprob, ntiles);
        else
            hipLaunchKernelGGL((gemm_batched_kernel<R, 2>), dim3((unsigned)ntiles), dim3(NT), 0, stream, d_probs,
                               d_tile_start, nprob, ntiles);
    };
    switch (var) {
    case 1:
        hipLaunchKernelGGL((gemm_batched_kernel<R, 1>), dim3((unsigned)ntiles), dim3(NT), 0, stream, d_probs,
                           d_tile_start, nprob, ntiles);
        break;
    case 2:
        tile64();
        break;
    case 3:
        hipLaunchKernelGGL((gemm_batched_kernel<R, 3>), dim3((unsigned)ntiles), dim3(NT), 0, stream, d_probs,
                           d_tile_start, nprob, ntiles);
        break;
    case WT_GEMM_SPLIT:
        if constexpr (sizeof(R) == 4)
            hipLaunchKernelGGL((gemm_split_kernel<128>), dim3((unsigned)ntiles), dim3(NT), 0, stream, d_probs,
                               d_tile_start, nprob, ntiles);
        else
            tile64();
        break;
    case WT_GEMM_SPLIT64:
        if constexpr (sizeof(R) == 4)
            hipLaunchKernelGGL((gemm_split_kernel<64>), dim3((unsigned)ntiles), dim3(NT), 0, stream, d_probs,
                               d_tile_start, nprob, ntiles);
        else
            tile64();
        break;
    default:
        hipLaunchKernelGGL((gemm_batched_kernel<R, 0>), dim3((unsigned)ntiles), dim3(NT), 0, stream, d_probs,
                           d_tile_start, nprob, ntiles);
        break;
    }
    WT_HIP(hipGetLastError());
    return WT_OK;
}

// One thread per (row, column) of every listed split product: the partials
// summed in their fixed order, then applied to C like the product itself.
template <typename R>
__global__ __launch_bounds__(256) void gemm_reduce_kernel(const wt_gemm_red<R>* __restrict__ reds,
                                                          const cpx<R>* __restrict__ ws)
{
    const wt_gemm_red<R> E = reds[blockIdx.z];
    const int n = blockIdx.y;
    const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= E.N || m >= E.M) return;
    const int64_t st = (int64_t)E.M * E.N;
    const cpx<R>* w = ws + E.woff + (int64_t)n * E.M + m;
    cpx<R> acc = w[0];
    for (int q = 1; q < E.S; ++q) acc = acc + w[q * st];
    cpx<R>* c = E.C + (int64_t)n * E.ldc + m;
    if (E.mode == WT_GEMM_SET) *c = acc;
    else if (E.mode == WT_GEMM_NSET) *c = cpx<R>(-acc.re, -acc.im);
    else if (E.mode == WT_GEMM_SUB) *c = *c - acc;
    else *c = *c + acc;
}

template <typename R>
int wt_gemm_reduce_launch(const wt_gemm_red<R>* d_reds, int nred, int64_t maxM, int64_t maxN, const cpx<R>* ws,
                          hipStream_t stream)
{
    if (nred <= 0 || maxM <= 0 || maxN <= 0) return WT_OK;
    if (nred > 65535 || maxN > 65535) return WT_ERR_ARG;
    hipLaunchKernelGGL(gemm_reduce_kernel<R>, dim3((unsigned)((maxM + 255) / 256), (unsigned)maxN, (unsigned)nred),
                       dim3(256), 0, stream, d_reds, ws);
    WT_HIP(hipGetLastError());
    return WT_OK;
}

template int wt_gemm_reduce_launch<float>(const wt_gemm_red<float>*, int, int64_t, int64_t, const cpx<float>*,
                                          hipStream_t);
template int wt_gemm_reduce_launch<double>(const wt_gemm_red<double>*, int, int64_t, int64_t, const cpx<double>*,
                                           hipStream_t);
template int wt_gemm_launch<float>(const wt_gemm_prob<float>*, const int64_t*, int, int64_t, int, hipStream_t, int);
template int wt_gemm_launch<double>(const wt_gemm_prob<double>*, const int64_t*, int, int64_t, int, hipStream_t, int);
