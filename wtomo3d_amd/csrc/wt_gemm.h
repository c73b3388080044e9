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

// Tile sizes of the kernel for precision R (rows x cols of C per block).
template <typename R>
struct wt_gemm_tile;
template <>
struct wt_gemm_tile<float> { static constexpr int BM = 128, BN = 128; };
template <>
struct wt_gemm_tile<double> { static constexpr int BM = 128, BN = 64; };

template <typename R>
inline int64_t wt_gemm_tiles(int64_t M, int64_t N)
{
    return ((M + wt_gemm_tile<R>::BM - 1) / wt_gemm_tile<R>::BM) *
           ((N + wt_gemm_tile<R>::BN - 1) / wt_gemm_tile<R>::BN);
}

// Launches nprob problems (device arrays probs[nprob], tile_start[nprob+1],
// tile_start[0] = 0) as one grid of tile_start[nprob] blocks.
template <typename R>
int wt_gemm_launch(const wt_gemm_prob<R>* d_probs, const int64_t* d_tile_start, int nprob,
                   int64_t ntiles, hipStream_t stream);
