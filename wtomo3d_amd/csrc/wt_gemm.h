// Batched complex GEMM on MFMA (3M form), used by every dense front kernel.
#pragma once
#include <cstdint>
#include <vector>
#include <hip/hip_runtime.h>

#include "wt_common.h"

enum wt_gemm_mode : int32_t {
    WT_GEMM_SUB = 0,  // C -= A B
    WT_GEMM_SET = 1,  // C  = A B   (C may alias B when one tile spans all M,
                      //             or alias A when one tile spans all N)
    WT_GEMM_ADD = 2,  // C += A B
    WT_GEMM_NSET = 3, // C = -A B
};

// One problem of a batch.  Column-major complex operands.
template <typename R>
struct wt_gemm_prob {
    const cpx<R>* A;
    const cpx<R>* B;
    cpx<R>* C;
    int64_t lda, ldb, ldc;
    int32_t M, N, K;
    int32_t mode;
    const int32_t* bperm;  // optional: row k of B is read from row bperm[k]
    int64_t pad;
};

// Tile shapes of the kernel for precision R: variant 0 is the square-ish
// default, variant 1 serves problems with N <= 64 (the block solves with
// 64 right-hand sides) without wasting half of each tile, variant 2 is a
// 64 x 64 tile for small fronts (more blocks per problem), variant 3 a
// 128 x 64 (fp64: 64 x 128) tile between them.  Every variant
// has BM, BN >= 64 = WT_NB, which the aliased panel products rely on (one
// tile spans all M, or all N, of a 64-wide panel).  The K loop is the same
// in variants 0-3, so they give identical results.  Variants 4 and 5 (fp32
// only) are 128 x 128 and 64 x 64 tiles whose real products run on the bf16
// matrix cores as three-way splits of each fp32 operand (wt_gemm.hip):
// results agree with variants 0-3 to fp32 rounding, not bit for bit.  For
// fp64, variants 4 and 5 are variant 2 again.
#define WT_GEMM_NVAR 6
#define WT_GEMM_SPLIT 4
#define WT_GEMM_SPLIT64 5
template <typename R, int V>
struct wt_gemm_tile;
template <>
struct wt_gemm_tile<float, 0> { static constexpr int BM = 128, BN = 128, WGM = 2, WGN = 2; };
template <>
struct wt_gemm_tile<float, 1> { static constexpr int BM = 256, BN = 64, WGM = 4, WGN = 1; };
template <>
struct wt_gemm_tile<float, 2> { static constexpr int BM = 64, BN = 64, WGM = 2, WGN = 2; };
template <>
struct wt_gemm_tile<double, 0> { static constexpr int BM = 128, BN = 64, WGM = 2, WGN = 2; };
template <>
struct wt_gemm_tile<double, 1> { static constexpr int BM = 128, BN = 64, WGM = 2, WGN = 2; };
template <>
struct wt_gemm_tile<double, 2> { static constexpr int BM = 64, BN = 64, WGM = 2, WGN = 2; };
template <>
struct wt_gemm_tile<float, 3> { static constexpr int BM = 128, BN = 64, WGM = 2, WGN = 2; };
template <>
struct wt_gemm_tile<double, 3> { static constexpr int BM = 64, BN = 128, WGM = 2, WGN = 2; };
template <>
struct wt_gemm_tile<float, 4> { static constexpr int BM = 128, BN = 128, WGM = 2, WGN = 2; };
template <>
struct wt_gemm_tile<double, 4> { static constexpr int BM = 64, BN = 64, WGM = 2, WGN = 2; };
template <>
struct wt_gemm_tile<float, 5> { static constexpr int BM = 64, BN = 64, WGM = 2, WGN = 2; };
template <>
struct wt_gemm_tile<double, 5> { static constexpr int BM = 64, BN = 64, WGM = 2, WGN = 2; };

// Variants worth measuring for precision R (fp64's variants 4 and 5 repeat 2).
template <typename R>
constexpr int wt_gemm_nvar()
{
    return sizeof(R) == 4 ? WT_GEMM_NVAR : WT_GEMM_SPLIT;
}

template <typename R>
inline int64_t wt_gemm_tiles(int64_t M, int64_t N, int var = 0)
{
    int BM, BN;
    switch (var) {
    case 1: BM = wt_gemm_tile<R, 1>::BM; BN = wt_gemm_tile<R, 1>::BN; break;
    case 2: BM = wt_gemm_tile<R, 2>::BM; BN = wt_gemm_tile<R, 2>::BN; break;
    case 3: BM = wt_gemm_tile<R, 3>::BM; BN = wt_gemm_tile<R, 3>::BN; break;
    case 4: BM = wt_gemm_tile<R, 4>::BM; BN = wt_gemm_tile<R, 4>::BN; break;
    case 5: BM = wt_gemm_tile<R, 5>::BM; BN = wt_gemm_tile<R, 5>::BN; break;
    default: BM = wt_gemm_tile<R, 0>::BM; BN = wt_gemm_tile<R, 0>::BN; break;
    }
    return ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
}

// A product split along K: S partial products, each written (mode SET,
// ld M) to ws[woff + s M N ...], are summed in order s = 0..S-1 and applied
// to C with the product's mode by wt_gemm_reduce_launch.
template <typename R>
struct wt_gemm_red {
    cpx<R>* C;
    int64_t ldc;
    int64_t woff;  // offset of the first partial in the workspace
    int32_t M, N, S, mode;
};

// Launches nred reductions (device array reds) over workspace ws; maxM and
// maxN bound the entries' M and N.
template <typename R>
int wt_gemm_reduce_launch(const wt_gemm_red<R>* d_reds, int nred, int64_t maxM, int64_t maxN, const cpx<R>* ws,
                          hipStream_t stream);

// Launches nprob problems (device arrays probs[nprob], tile_start[nprob+1],
// tile_start[0] = 0) as one grid of tile_start[nprob] blocks of variant var.
// rmw != 0 declares that every problem reads C (WT_GEMM_SUB / WT_GEMM_ADD):
// the 64 x 64 tile then loads C ahead, with identical results.
template <typename R>
int wt_gemm_launch(const wt_gemm_prob<R>* d_probs, const int64_t* d_tile_start, int nprob,
                   int64_t ntiles, int var, hipStream_t stream, int rmw = 0);
