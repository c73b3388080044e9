// Multifrontal solver state: symbolic analysis, device arenas, and the
// launch plans of the numeric factorization and the block solve.
#pragma once
#include <cstdint>
#include <utility>
#include <vector>
#include <hip/hip_runtime.h>

#include "wt_front.h"
#include "wt_gemm.h"
#include "wt_symbolic.h"

// One launch of the batched GEMM: a contiguous range of the plan's problems.
// Tile counts are kept for every tile variant; var is the variant used (the
// default rule of end(), or the autotuner's measured choice).
struct wt_gemm_launch_rec {
    int64_t prob_off, tile_off;
    int64_t ntiles[WT_GEMM_NVAR];
    int32_t nprob;
    int32_t var;
    double flops;  // 8 M N K summed over the launch's problems
    // products split along K: their reductions, run after the launch
    int64_t red_off, ws_len, red_maxM, red_maxN;
    int32_t nred;
    int32_t rmw;  // every problem reads C (WT_GEMM_SUB / WT_GEMM_ADD)
};

template <typename R>
struct wt_gemm_plan {
    std::vector<wt_gemm_prob<R>> probs;
    std::vector<int64_t> tiles[WT_GEMM_NVAR];  // per-variant prefix sums, same indexing
    std::vector<wt_gemm_launch_rec> launches;
    wt_gemm_prob<R>* d_probs = nullptr;
    int64_t* d_tiles = nullptr;  // WT_GEMM_NVAR arrays of tiles[0].size(), variant-major
    int64_t tstride = 0;
    wt_gemm_launch_rec cur{};
    bool allow_split = true;  // fp32: the bf16-split tile may be chosen
    // split-K products: reductions, and the partials' workspace (each
    // launch's partials start at offset 0: launches run in stream order and
    // a launch's reduction consumes them before the next launch)
    std::vector<wt_gemm_red<R>> reds;
    std::vector<std::pair<int64_t, int64_t>> ws_patch;  // (problem, offset): C = ws + offset at upload
    wt_gemm_red<R>* d_reds = nullptr;
    cpx<R>* ws = nullptr;
    int64_t ws_cap = 0;
    void begin()
    {
        cur.prob_off = (int64_t)probs.size();
        cur.tile_off = (int64_t)tiles[0].size();
        for (int v = 0; v < WT_GEMM_NVAR; ++v) {
            cur.ntiles[v] = 0;
            tiles[v].push_back(0);
        }
        cur.nprob = 0;
        cur.flops = 0;
        cur.red_off = (int64_t)reds.size();
        cur.nred = 0;
        cur.ws_len = cur.red_maxM = cur.red_maxN = 0;
    }
    // C (mode) A B as S products over consecutive K ranges (multiples of 16
    // deep) whose partials a reduction after the launch sums in order and
    // applies: more blocks for a product with few output tiles and a long K.
    // S <= 1 adds the product whole.
    void add_split(const cpx<R>* A, int64_t lda, const cpx<R>* B, int64_t ldb, cpx<R>* C, int64_t ldc, int64_t M,
                   int64_t N, int64_t K, int mode, int S)
    {
        if (M <= 0 || N <= 0 || K <= 0) return;
        const int64_t chunk = S > 1 ? ((K + S - 1) / S + 15) / 16 * 16 : K;
        if (S <= 1 || chunk >= K) {
            add(A, lda, B, ldb, C, ldc, M, N, K, mode);
            return;
        }
        wt_gemm_red<R> r;
        r.C = C; r.ldc = ldc; r.woff = cur.ws_len; r.M = (int32_t)M; r.N = (int32_t)N; r.mode = mode;
        int q = 0;
        for (int64_t k0 = 0; k0 < K; k0 += chunk, ++q) {
            const int64_t kk = K - k0 < chunk ? K - k0 : chunk;
            ws_patch.push_back({(int64_t)probs.size(), cur.ws_len});
            add(A + k0 * lda, lda, B + k0, ldb, nullptr, M, M, N, kk, WT_GEMM_SET);
            cur.ws_len += M * N;
        }
        r.S = q;
        reds.push_back(r);
        ++cur.nred;
        cur.red_maxM = M > cur.red_maxM ? M : cur.red_maxM;
        cur.red_maxN = N > cur.red_maxN ? N : cur.red_maxN;
    }
    void add(const cpx<R>* A, int64_t lda, const cpx<R>* B, int64_t ldb, cpx<R>* C, int64_t ldc, int64_t M,
             int64_t N, int64_t K, int mode, const int32_t* bperm = nullptr)
    {
        if (M <= 0 || N <= 0 || K <= 0) return;
        wt_gemm_prob<R> p;
        p.A = A; p.B = B; p.C = C; p.lda = lda; p.ldb = ldb; p.ldc = ldc;
        p.M = (int32_t)M; p.N = (int32_t)N; p.K = (int32_t)K; p.mode = mode; p.bperm = bperm; p.pad = 0;
        probs.push_back(p);
        for (int v = 0; v < WT_GEMM_NVAR; ++v) {
            cur.ntiles[v] += wt_gemm_tiles<R>(M, N, v);
            tiles[v].push_back(cur.ntiles[v]);
        }
        cur.flops += 8.0 * (double)M * (double)N * (double)K;
        ++cur.nprob;
    }
    // returns the launch index, or -1 when nothing was added
    int end()
    {
        if (cur.nprob == 0) {
            for (int v = 0; v < WT_GEMM_NVAR; ++v) tiles[v].pop_back();
            return -1;
        }
        // default tile: 64 x 64 unless the launch has enough large tiles to
        // fill the GPU several times over and they pad the problems little
        // more than 64 x 64 tiles do (measured: the large tile wins only on
        // the largest updates; narrow, small and ragged launches run up to
        // 2x faster on 64 x 64, which gives them 4x the blocks)
        bool narrow = true;
        for (int64_t i = cur.prob_off; i < (int64_t)probs.size(); ++i)
            if (probs[i].N > 64) narrow = false;
        const double a0 = (double)cur.ntiles[0] * wt_gemm_tile<R, 0>::BM * wt_gemm_tile<R, 0>::BN;
        const double a2 = (double)cur.ntiles[2] * wt_gemm_tile<R, 2>::BM * wt_gemm_tile<R, 2>::BN;
        cur.var = (!narrow && cur.ntiles[0] >= 1024 && a0 <= 1.15 * a2) ? 0 : 2;
        // fp32 large launches: the same 128 x 128 tile with its products as
        // bf16 splits, unless the handle keeps to the fp32 matrix cores
        if (sizeof(R) == 4 && cur.var == 0 && allow_split) cur.var = WT_GEMM_SPLIT;
        // fp64: the 64 x 64 tile throughout (two blocks per CU; measured
        // fastest on large and narrow launches alike)
        if (sizeof(R) == 8) cur.var = 2;
        cur.rmw = 1;
        for (int64_t i = cur.prob_off; i < (int64_t)probs.size(); ++i)
            if (probs[i].mode != WT_GEMM_SUB && probs[i].mode != WT_GEMM_ADD) cur.rmw = 0;
        launches.push_back(cur);
        return (int)launches.size() - 1;
    }
    int upload();
    void release();
    int run(int i, hipStream_t s) const
    {
        const wt_gemm_launch_rec& L = launches[i];
        int rc = wt_gemm_launch<R>(d_probs + L.prob_off, d_tiles + L.var * tstride + L.tile_off, L.nprob,
                                   L.ntiles[L.var], L.var, s, L.rmw);
        if (rc == WT_OK && L.nred > 0)
            rc = wt_gemm_reduce_launch<R>(d_reds + L.red_off, L.nred, L.red_maxM, L.red_maxN, ws, s);
        return rc;
    }
};

enum wt_step_kind {
    WT_STEP_ZERO, WT_STEP_EXTADD, WT_STEP_DIAG, WT_STEP_GEMM, WT_STEP_SGATHER, WT_STEP_SGATHER_UPD,
    WT_STEP_SPFWD, WT_STEP_SPBWD
};

struct wt_step {
    int kind;
    int32_t count;       // list length (diag / extend-add / gathers)
    int64_t a;           // ZERO: entries; DIAG: kb; EXTADD/SGATHER: max_u
    int64_t list_off;    // offset into the device list array
    int32_t buf;         // arena parity (ZERO: which arena; EXTADD: parent parity)
    int32_t gemm;        // GEMM launch index
    int32_t part = 0;    // EXTADD: 1 = into the parents' pivot blocks, 2 = into their CBs
};

template <typename R>
struct wt_numeric {
    bool ready = false;
    cpx<R>* factor = nullptr;
    cpx<R>* cb[2] = {nullptr, nullptr};
    cpx<R>* tiles = nullptr;
    int32_t* perm = nullptr;
    int* status = nullptr;  // WT_PIVSTAT_LEN pivot statistics on the device
    int pivstat[WT_PIVSTAT_LEN] = {0, 0, 0x7f800000, 0};  // host copy after the factorization
    bool have_pivstat = false;
    wt_gemm_plan<R> fplan;
    std::vector<wt_step> fsteps;
    // block solve
    int snrhs = 0;
    cpx<R>* y = nullptr;
    cpx<R>* w[2] = {nullptr, nullptr};
    wt_gemm_plan<R> splan;
    std::vector<wt_step> fwd, bwd;
    std::vector<int32_t> hsp;   // super-panel entries {front, K0, K1}
    int32_t* d_sp = nullptr;
    // captured sequences: the factorization, and one refinement step of a
    // batch solve (for sgraph_nrhs columns)
    hipGraphExec_t fgraph = nullptr;
    hipGraphExec_t sgraph = nullptr;
    int sgraph_nrhs = 0;
    bool tuned = false;  // GEMM tile variants chosen by measurement
    // factor, cb, tiles, perm, status, y and w are the caller's buffers
    // (a pool shared by the groups of a batch run): never allocated or freed
    // here, and y, w must hold the solve blocks of the largest nrhs used
    bool ext = false;
    void drop_graphs();
    void release();
};

// Per-kernel-class timing with HIP events on the solver stream (enabled by
// wt_solver_profile): launches, summed kernel time and algorithmic work.
enum wt_prof_class {
    WT_PROF_GEMM_FACTOR, WT_PROF_GEMM_SOLVE, WT_PROF_DIAG, WT_PROF_EXTADD, WT_PROF_FILL,
    WT_PROF_ASSEMBLE, WT_PROF_RESIDUAL, WT_PROF_PERMUTE, WT_PROF_TRSM, WT_PROF_GEMM_FACTOR_NARROW,
    WT_PROF_GEMM_FACTOR_SPLIT, WT_PROF_NCLASS
};
struct wt_prof {
    bool on = false;
    struct pend { int cls; double flops, bytes; hipEvent_t a, b; };
    std::vector<pend> pending;
    std::vector<hipEvent_t> pool;
    int64_t launches[WT_PROF_NCLASS] = {};
    double ms[WT_PROF_NCLASS] = {}, flops[WT_PROF_NCLASS] = {}, bytes[WT_PROF_NCLASS] = {};
    hipEvent_t take()
    {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    // records an event pair around fn() when enabled
    template <typename F>
    int run(hipStream_t s, int cls, double fl, double by, F&& fn)
    {
        if (!on) return fn();
        hipEvent_t a = take(), b = take();
        (void)hipEventRecord(a, s);
        int rc = fn();
        (void)hipEventRecord(b, s);
        pending.push_back({cls, fl, by, a, b});
        return rc;
    }
    // after the stream has been synchronized
    void flush()
    {
        for (auto& p : pending) {
            float t = 0;
            if (hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) ms[p.cls] += t;
            launches[p.cls]++;
            flops[p.cls] += p.flops;
            bytes[p.cls] += p.bytes;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    void reset()
    {
        pending.clear();
        for (int i = 0; i < WT_PROF_NCLASS; ++i) { launches[i] = 0; ms[i] = flops[i] = bytes[i] = 0; }
    }
    void release()
    {
        for (auto& p : pending) { pool.push_back(p.a); pool.push_back(p.b); }
        pending.clear();
        for (auto e : pool) (void)hipEventDestroy(e);
        pool.clear();
    }
};

struct wt_solver_impl {
    wt_prof prof;
    wt_symbolic S;
    hipStream_t stream = nullptr;
    int64_t n = 0, nnz = 0;
    // device copies of the symbolic data
    int32_t* rowptr = nullptr;
    int32_t* colidx = nullptr;
    int64_t* dst = nullptr;
    int64_t* iperm = nullptr;
    int64_t* perm = nullptr;  // perm[iperm[i]] = i
    int64_t* upd = nullptr;
    int32_t* relmap = nullptr;
    wt_dfront* fronts = nullptr;
    int32_t* lists = nullptr;
    std::vector<int32_t> hlists;
    // beside every child-list entry of hlists: its first extend-add block
    // (the list's exclusive prefix sum of ceil(u / WT_EA_COLS))
    int32_t* cols = nullptr;
    std::vector<int32_t> hcols;
    // per level: fronts sorted by p (desc), and children by slot
    std::vector<int64_t> lev_sorted_off;
    std::vector<std::vector<int64_t>> lev_sorted_p;
    std::vector<int64_t> lev_kids_off[2];
    std::vector<int32_t> lev_kids_cnt[2];
    std::vector<int64_t> lev_kids_maxu[2];
    std::vector<int64_t> lev_kids_blocks[2];  // extend-add blocks of the list
    std::vector<int32_t> lev_upd_cnt;
    std::vector<int64_t> lev_upd_off;
    // model and matrix
    cpx<double>* lam = nullptr;
    cpx<double>* mu = nullptr;
    double* rho = nullptr;
    cpx<double>* vals = nullptr;  // CSR values, complex128
    double* rscale = nullptr;     // row equilibration of the factored matrix
    bool have_vals = false;
    wt_grid_params last_p{};  // parameters of the last assembly (have_p)
    bool have_p = false;
    int* d_err = nullptr;
    // refinement buffers (row-major n x 64)
    int ir_cap = 0;
    cpx<double>* xb = nullptr;
    cpx<double>* rb = nullptr;
    cpx<double>* bb = nullptr;
    double* partial = nullptr;
    double* norms = nullptr;
    wt_numeric<float> f32;
    wt_numeric<double> f64;
    int factored_prec = 0;
    double last_factor_ms = 0;
    bool use_graphs = false;  // capture and replay the factorization
    bool gemm_split = true;   // fp32 GEMMs may run as bf16 splits (variant WT_GEMM_SPLIT)
    int splitk_depth = 512;   // least K of a piece of a split back-substitution product (0: no splits)
};

// Row tiles a back-substitution update launch is split along K to reach.
#define WT_SPLITK_BLOCKS 2048

struct wt_grid_dims {
    int nx, ny, nz;
};

int wt_solver_init(wt_solver_impl& s, int nx, int ny, int nz, int leaf_nodes);
int wt_solver_init_union(wt_solver_impl& s, const std::vector<wt_grid_dims>& grids, int leaf_nodes);
int wt_union_build(const std::vector<wt_grid_dims>& grids, int leaf_nodes, wt_symbolic& U,
                   std::vector<int32_t>* hr, std::vector<int32_t>* hc);
void wt_solver_free(wt_solver_impl& s);
template <typename R>
int wt_solver_factor_run(wt_solver_impl& s, wt_numeric<R>& num);
template <typename R>
int wt_solver_factor_launch(wt_solver_impl& s, wt_numeric<R>& num);
template <typename R>
int wt_solver_factor_wait(wt_solver_impl& s, wt_numeric<R>& num);
template <typename R>
int wt_solver_autotune(wt_solver_impl& s, wt_numeric<R>& num, double* speedup);
template <typename R>
int wt_solver_plan_tiles(wt_solver_impl& s, wt_numeric<R>& num, int which, int nrhs, int32_t* vars,
                         const int32_t* in, int64_t* count);
// Refinement buffers (xb, rb, bb, partial, norms) for nrhs columns.
int wt_solver_ir_buffers(wt_solver_impl& s, int nrhs);
template <typename R>
int wt_solver_prepare_solve(wt_solver_impl& s, wt_numeric<R>& num, int nrhs);
template <typename R>
int wt_solver_solve_block(wt_solver_impl& s, wt_numeric<R>& num, int nrhs, cpx<R>* y);
