// Front-level kernels of the multifrontal factorization and solve.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

#include "wt_common.h"

// Diagonal tile width: pivoting is restricted to rows of one tile.
#define WT_NB 64
// Super-panel width of the two-level blocked factorization of each pivot
// block: trailing updates beyond the super-panel are deferred to one
// rank-WT_NB2 GEMM (a multiple of WT_NB).
#define WT_NB2 512
// Super-panel width of the factorization alone (a multiple of WT_NB): the
// rank-WT_NB updates inside a super-panel have K = 64, the deferred update
// has K = WT_FNB2.  Independent of the solve's WT_NB2: the layout of the
// stored factor (packed L\U and 64 x 64 tile inverses) does not depend on it.
#ifndef WT_FNB2
#define WT_FNB2 512
#endif

// Device copy of the per-front offsets of wt_front (wt_symbolic.h).
struct wt_dfront {
    int64_t p, u, pstart;
    int64_t Loff, Uoff, CBoff, moff, uoff, tileoff, pivoff, Woff;
    int32_t parent, level, slot, pad;
};

int wt_zero_i32(int* p, int n, hipStream_t s);
// Pivot statistics of one factorization, kept by the diagonal-tile kernels
// in WT_PIVSTAT_LEN ints: [0] 1 if an exactly zero pivot was replaced,
// [1] the number of pivots with |d| <= sqrt(eps) max|tile|, [2] the bits of
// the smallest |d| / max|tile|, [3] the bits of the largest max|U| / max|tile|
// (non-negative floats, whose bits order as unsigned integers).  Magnitudes
// are |re| + |im|; max|tile| is taken when the tile's elimination starts.
#define WT_PIVSTAT_LEN 4
int wt_pivstat_init(int* p, hipStream_t s);
template <typename R>
int wt_fill_zero(cpx<R>* p, int64_t n, hipStream_t s);
template <typename Ri, typename R>
int wt_scatter(const cpx<Ri>* vals, const int64_t* dst, int64_t nnz, cpx<R>* arena, hipStream_t s);
int wt_row_scale(const int32_t* rowptr, const cpx<double>* vals, int64_t n, double* scale, hipStream_t s);
template <typename R>
int wt_scatter_rows(const int32_t* rowptr, const cpx<double>* vals, const int64_t* dst, const double* scale,
                    int64_t n, cpx<R>* arena, hipStream_t s);
// Contribution-block columns per extend-add block.  A launch covers a list of
// children whose column blocks are numbered consecutively: col_start[i] is
// the first block of child i (exclusive prefix sum of ceil(u / WT_EA_COLS)
// over the list) and nblocks the total, so no block is launched empty.
constexpr int WT_EA_COLS = 8;
template <typename R>
int wt_extend_add(const wt_dfront* fr, const int32_t* kids, const int32_t* col_start, int nkids, int64_t nblocks,
                  const int32_t* relmap, const cpx<R>* cb_child, cpx<R>* factor, cpx<R>* cb_parent, int part,
                  hipStream_t s);
template <typename R>
int wt_diag_tiles(const wt_dfront* fr, const int32_t* list, int nlist, int64_t kb, cpx<R>* factor, cpx<R>* tiles,
                  int32_t* perm, int* status, hipStream_t s);
template <typename R>
int wt_solve_gather(const wt_dfront* fr, const int32_t* kids, int nkids, const int32_t* relmap,
                    const cpx<R>* w_child, cpx<R>* w_parent, cpx<R>* y, int64_t n, int nrhs, hipStream_t s);
template <typename R>
int wt_solve_gather_upd(const wt_dfront* fr, const int32_t* list, int nlist, const int64_t* upd_idx,
                        const cpx<R>* x, cpx<R>* work, int64_t n, int nrhs, hipStream_t s);
template <typename R>
int wt_sp_solve(int fwd, const wt_dfront* fr, const int32_t* ent, int nent, const cpx<R>* factor,
                const cpx<R>* tiles, const int32_t* perm, cpx<R>* y, int64_t n, int nrhs, hipStream_t s);
