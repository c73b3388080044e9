// Host orchestration of the multifrontal factorization and block solve.
//
// The elimination tree is processed one level at a time, bottom-up for the
// factorization and the forward substitution, top-down for the back
// substitution.  All fronts of a level advance together through their
// nb-wide pivot tiles: one launch factors the diagonal tiles of every front,
// one batched-GEMM launch forms the row and column panels, one more applies
// the trailing update inside the pivot block, and a final launch per level
// forms every contribution block CB = -L21 U12 (the large-K Schur GEMM).
// Launch lists are built once per factorization buffer set and replayed.
#include "wt_solver.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include "wt_rt.h"
#include "wt_vec.h"

int wt_launch_pattern(int nx, int ny, int nz, int base, int32_t* d_rowptr, int32_t* d_colidx, hipStream_t stream);
int64_t wt_nnz(int nx, int ny, int nz);

template <typename T>
static int dalloc(T** p, int64_t count)
{
    *p = nullptr;
    if (count <= 0) count = 1;
    if (rt_malloc((void**)p, (size_t)count * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return WT_ERR_NOMEM;
    }
    return WT_OK;
}
template <typename T>
static void dfree(T*& p)
{
    if (p) (void)rt_free((void*)p);
    p = nullptr;
}

template <typename R>
int wt_gemm_plan<R>::upload()
{
    dfree(d_probs);
    dfree(d_tiles);
    dfree(d_reds);
    tstride = (int64_t)tiles[0].size();
    if (dalloc(&d_probs, (int64_t)probs.size()) || dalloc(&d_tiles, WT_GEMM_NVAR * tstride)) return WT_ERR_NOMEM;
    // the split products' partials: one workspace for the largest launch
    int64_t need = 0;
    for (const auto& L : launches) need = L.ws_len > need ? L.ws_len : need;
    if (need > ws_cap) {
        dfree(ws);
        ws_cap = 0;
        if (dalloc(&ws, need)) return WT_ERR_NOMEM;
        ws_cap = need;
    }
    for (const auto& pt : ws_patch) probs[pt.first].C = ws + pt.second;
    if (!reds.empty()) {
        if (dalloc(&d_reds, (int64_t)reds.size())) return WT_ERR_NOMEM;
        WT_HIP(rt_memcpy(d_reds, reds.data(), reds.size() * sizeof(reds[0]), hipMemcpyHostToDevice));
    }
    if (!probs.empty())
        WT_HIP(rt_memcpy(d_probs, probs.data(), probs.size() * sizeof(probs[0]), hipMemcpyHostToDevice));
    for (int v = 0; v < WT_GEMM_NVAR && tstride > 0; ++v)
        WT_HIP(rt_memcpy(d_tiles + v * tstride, tiles[v].data(), tstride * sizeof(int64_t), hipMemcpyHostToDevice));
    return WT_OK;
}
template <typename R>
void wt_gemm_plan<R>::release()
{
    dfree(d_probs);
    dfree(d_tiles);
    dfree(d_reds);
    dfree(ws);
    ws_cap = 0;
    probs.clear();
    for (int v = 0; v < WT_GEMM_NVAR; ++v) tiles[v].clear();
    launches.clear();
    reds.clear();
    ws_patch.clear();
    tstride = 0;
}

template <typename R>
void wt_numeric<R>::drop_graphs()
{
    if (fgraph) (void)hipGraphExecDestroy(fgraph);
    fgraph = nullptr;
    if (sgraph) (void)hipGraphExecDestroy(sgraph);
    sgraph = nullptr;
    sgraph_nrhs = 0;
}

template <typename R>
void wt_numeric<R>::release()
{
    drop_graphs();
    tuned = false;
    if (ext) {
        factor = cb[0] = cb[1] = tiles = y = w[0] = w[1] = nullptr;
        perm = nullptr;
        status = nullptr;
        ext = false;
    }
    dfree(factor);
    dfree(cb[0]);
    dfree(cb[1]);
    dfree(tiles);
    dfree(perm);
    dfree(status);
    dfree(y);
    dfree(w[0]);
    dfree(w[1]);
    dfree(d_sp);
    hsp.clear();
    fplan.release();
    splan.release();
    fsteps.clear();
    fwd.clear();
    bwd.clear();
    snrhs = 0;
    ready = false;
}

static int upload_symbolic(wt_solver_impl& s);

int wt_solver_init(wt_solver_impl& s, int nx, int ny, int nz, int leaf_nodes)
{
    if (nx < 2 || ny < 2 || nz < 2) return WT_ERR_ARG;
    WT_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    s.n = 3LL * nx * ny * nz;
    s.nnz = wt_nnz(nx, ny, nz);
    if (s.nnz >= (1LL << 31)) return WT_ERR_ARG;
    int rc;
    if ((rc = dalloc(&s.rowptr, s.n + 1)) || (rc = dalloc(&s.colidx, s.nnz))) return rc;
    if ((rc = wt_launch_pattern(nx, ny, nz, 0, s.rowptr, s.colidx, s.stream))) return rc;
    std::vector<int32_t> hr(s.n + 1), hc(s.nnz);
    WT_HIP(hipMemcpyAsync(hr.data(), s.rowptr, (s.n + 1) * 4, hipMemcpyDeviceToHost, s.stream));
    WT_HIP(hipMemcpyAsync(hc.data(), s.colidx, s.nnz * 4, hipMemcpyDeviceToHost, s.stream));
    WT_HIP(hipStreamSynchronize(s.stream));
    if ((rc = wt_symbolic_build(nx, ny, nz, leaf_nodes, WT_NB, hr.data(), hc.data(), s.S))) return rc;
    hc.clear();
    hc.shrink_to_fit();
    return upload_symbolic(s);
}

// Independent grids as one block-diagonal system (a forest of elimination
// trees advancing level by level together): per-grid patterns and symbolic
// analyses on the host, in parallel, then their union.  hr/hc receive the
// union's CSR pattern (0-based) when given.
int wt_union_build(const std::vector<wt_grid_dims>& grids, int leaf_nodes, wt_symbolic& U,
                   std::vector<int32_t>* hr, std::vector<int32_t>* hc)
{
    const int m = (int)grids.size();
    if (m < 1) return WT_ERR_ARG;
    std::vector<int64_t> noff(m + 1, 0), eoff(m + 1, 0);
    for (int k = 0; k < m; ++k) {
        const wt_grid_dims& g = grids[k];
        if (g.nx < 2 || g.ny < 2 || g.nz < 2) return WT_ERR_ARG;
        noff[k + 1] = noff[k] + 3LL * g.nx * g.ny * g.nz;
        eoff[k + 1] = eoff[k] + wt_nnz(g.nx, g.ny, g.nz);
    }
    if (eoff[m] >= (1LL << 31)) return WT_ERR_ARG;
    if (hr) hr->assign(noff[m] + 1, 0);
    if (hc) hc->assign(eoff[m], 0);
    std::vector<wt_symbolic> parts(m);
    std::vector<int> res(m, WT_OK);
    {
        const unsigned nth = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; ++t)
            th.emplace_back([&, t] {
                std::vector<int32_t> r, c;
                for (int k = (int)t; k < m; k += (int)nth) {
                    const wt_grid_dims& g = grids[k];
                    const int64_t n = noff[k + 1] - noff[k], e = eoff[k + 1] - eoff[k];
                    r.resize(n + 1);
                    c.resize(e);
                    int rc = wt_pattern_host(g.nx, g.ny, g.nz, 0, 0, r.data(), c.data());
                    if (!rc) rc = wt_symbolic_build(g.nx, g.ny, g.nz, leaf_nodes, WT_NB, r.data(), c.data(), parts[k]);
                    res[k] = rc;
                    if (rc) continue;
                    if (hr)
                        for (int64_t i = 0; i < n; ++i) (*hr)[noff[k] + i] = (int32_t)(r[i] + eoff[k]);
                    if (hc)
                        for (int64_t i = 0; i < e; ++i) (*hc)[eoff[k] + i] = (int32_t)(c[i] + noff[k]);
                }
            });
        for (auto& x : th) x.join();
    }
    for (int k = 0; k < m; ++k)
        if (res[k]) return res[k];
    if (hr) (*hr)[noff[m]] = (int32_t)eoff[m];
    std::vector<const wt_symbolic*> pp(m);
    for (int k = 0; k < m; ++k) pp[k] = &parts[k];
    return wt_symbolic_union(pp, U);
}

int wt_solver_init_union(wt_solver_impl& s, const std::vector<wt_grid_dims>& grids, int leaf_nodes)
{
    std::vector<int32_t> hr, hc;
    int rc = wt_union_build(grids, leaf_nodes, s.S, &hr, &hc);
    if (rc) return rc;
    WT_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    s.n = s.S.n;
    s.nnz = s.S.nnz;
    if ((rc = dalloc(&s.rowptr, s.n + 1)) || (rc = dalloc(&s.colidx, s.nnz))) return rc;
    WT_HIP(rt_memcpy(s.rowptr, hr.data(), (s.n + 1) * 4, hipMemcpyHostToDevice));
    WT_HIP(rt_memcpy(s.colidx, hc.data(), s.nnz * 4, hipMemcpyHostToDevice));
    return upload_symbolic(s);
}

// Device copies of the symbolic data (front table, per-level lists, CSR
// scatter map, permutation, update sets, child->parent maps).
static int upload_symbolic(wt_solver_impl& s)
{
    int rc;
    const wt_symbolic& S = s.S;

    // device front table
    std::vector<wt_dfront> df(S.fronts.size());
    for (size_t f = 0; f < S.fronts.size(); ++f) {
        const wt_front& F = S.fronts[f];
        wt_dfront& d = df[f];
        d.p = F.p; d.u = F.u; d.pstart = F.pstart; d.Loff = F.Loff; d.Uoff = F.Uoff; d.CBoff = F.CBoff;
        d.moff = F.moff; d.uoff = F.uoff; d.tileoff = F.tileoff; d.pivoff = F.pivoff; d.Woff = F.Woff;
        d.parent = F.parent; d.level = F.level; d.slot = F.slot; d.pad = 0;
    }
    // per-level lists
    const int nl = (int)S.levels.size();
    s.lev_sorted_off.assign(nl, 0);
    s.lev_sorted_p.assign(nl, {});
    for (int k = 0; k < 2; ++k) {
        s.lev_kids_off[k].assign(nl, 0);
        s.lev_kids_cnt[k].assign(nl, 0);
        s.lev_kids_maxu[k].assign(nl, 0);
        s.lev_kids_blocks[k].assign(nl, 0);
    }
    s.lev_upd_cnt.assign(nl, 0);
    s.lev_upd_off.assign(nl, 0);
    s.hlists.clear();
    s.hcols.clear();
    for (int l = 0; l < nl; ++l) {
        std::vector<int32_t> v = S.levels[l];
        std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return S.fronts[a].p > S.fronts[b].p; });
        s.lev_sorted_off[l] = (int64_t)s.hlists.size();
        for (int f : v) {
            s.hlists.push_back(f);
            s.lev_sorted_p[l].push_back(S.fronts[f].p);
        }
        for (int k = 0; k < 2; ++k) {
            s.lev_kids_off[k][l] = (int64_t)s.hlists.size();
            s.hcols.resize(s.hlists.size(), 0);
            int64_t nb = 0;
            for (int f : S.levels[l])
                if (S.fronts[f].parent >= 0 && S.fronts[f].slot == k && S.fronts[f].u > 0) {
                    s.hlists.push_back(f);
                    s.hcols.push_back((int32_t)nb);
                    nb += (S.fronts[f].u + WT_EA_COLS - 1) / WT_EA_COLS;
                    s.lev_kids_cnt[k][l]++;
                    s.lev_kids_maxu[k][l] = std::max(s.lev_kids_maxu[k][l], S.fronts[f].u);
                }
            if (nb > 0x7fffffff) return WT_ERR_ARG;
            s.lev_kids_blocks[k][l] = nb;
        }
        s.lev_upd_off[l] = (int64_t)s.hlists.size();
        for (int f : S.levels[l])
            if (S.fronts[f].u > 0) {
                s.hlists.push_back(f);
                s.lev_upd_cnt[l]++;
            }
    }
    s.hcols.resize(s.hlists.size(), 0);
    if ((rc = dalloc(&s.fronts, (int64_t)df.size())) || (rc = dalloc(&s.lists, (int64_t)s.hlists.size())) ||
        (rc = dalloc(&s.cols, (int64_t)s.hcols.size())) ||
        (rc = dalloc(&s.dst, S.nnz)) || (rc = dalloc(&s.iperm, S.n)) || (rc = dalloc(&s.perm, S.n)) || (rc = dalloc(&s.upd, (int64_t)S.upd_idx.size())) ||
        (rc = dalloc(&s.relmap, (int64_t)S.relmap.size())) || (rc = dalloc(&s.d_err, 1)))
        return rc;
    WT_HIP(rt_memcpy(s.fronts, df.data(), df.size() * sizeof(wt_dfront), hipMemcpyHostToDevice));
    WT_HIP(rt_memcpy(s.lists, s.hlists.data(), s.hlists.size() * 4, hipMemcpyHostToDevice));
    WT_HIP(rt_memcpy(s.cols, s.hcols.data(), s.hcols.size() * 4, hipMemcpyHostToDevice));
    WT_HIP(rt_memcpy(s.dst, S.dst.data(), S.nnz * 8, hipMemcpyHostToDevice));
    WT_HIP(rt_memcpy(s.iperm, S.iperm.data(), S.n * 8, hipMemcpyHostToDevice));
    WT_HIP(rt_memcpy(s.perm, S.perm.data(), S.n * 8, hipMemcpyHostToDevice));
    if (!S.upd_idx.empty())
        WT_HIP(rt_memcpy(s.upd, S.upd_idx.data(), S.upd_idx.size() * 8, hipMemcpyHostToDevice));
    if (!S.relmap.empty())
        WT_HIP(rt_memcpy(s.relmap, S.relmap.data(), S.relmap.size() * 4, hipMemcpyHostToDevice));
    // the dst map now lives on the device only
    s.S.dst.clear();
    s.S.dst.shrink_to_fit();
    return WT_OK;
}

void wt_solver_free(wt_solver_impl& s)
{
    s.prof.release();
    s.f32.release();
    s.f64.release();
    dfree(s.rowptr); dfree(s.colidx); dfree(s.dst); dfree(s.iperm); dfree(s.perm); dfree(s.upd); dfree(s.relmap);
    dfree(s.fronts); dfree(s.lists); dfree(s.cols); dfree(s.rscale); dfree(s.lam); dfree(s.mu); dfree(s.rho); dfree(s.vals); dfree(s.d_err);
    dfree(s.xb); dfree(s.rb); dfree(s.bb); dfree(s.partial); dfree(s.norms);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    s.stream = nullptr;
}

// number of fronts of level l with p > kb (lists are sorted by p, desc)
static int count_gt(const std::vector<int64_t>& ps, int64_t kb)
{
    int c = 0;
    while (c < (int)ps.size() && ps[c] > kb) ++c;
    return c;
}

template <typename R>
static int build_factor_plan(wt_solver_impl& s, wt_numeric<R>& num)
{
    const wt_symbolic& S = s.S;
    const int nl = (int)S.levels.size();
    const int64_t nb = S.nb;
    const int64_t NB2 = WT_FNB2;
    num.fplan.release();
    num.fplan.allow_split = s.gemm_split;
    num.fsteps.clear();
    // The contribution blocks need no clearing: each front's CB is first
    // written by its own CB = -L21 U12 product, and its children's
    // contributions are split around it (into the pivot blocks before the
    // front is factored, into the CB after the product).
    auto extend_add = [&](int l, int part) {
        if (l + 1 >= nl) return;
        for (int k = 0; k < 2; ++k)
            if (s.lev_kids_cnt[k][l + 1] > 0) {
                wt_step t{WT_STEP_EXTADD, s.lev_kids_cnt[k][l + 1], s.lev_kids_blocks[k][l + 1],
                          s.lev_kids_off[k][l + 1], l & 1, -1};
                t.part = part;
                num.fsteps.push_back(t);
            }
    };
    for (int l = nl - 1; l >= 0; --l) {
        const int par = l & 1;
        extend_add(l, 1);
        const std::vector<int64_t>& ps = s.lev_sorted_p[l];
        const int32_t* lst = s.hlists.data() + s.lev_sorted_off[l];
        for (int64_t kb = 0; kb < ps[0]; kb += nb) {
            const int cnt = count_gt(ps, kb);
            num.fsteps.push_back({WT_STEP_DIAG, cnt, kb, s.lev_sorted_off[l], par, -1});
            num.fplan.begin();
            for (int i = 0; i < cnt; ++i) {
                const wt_front& F = S.fronts[lst[i]];
                const int64_t p = F.p, u = F.u, ld = p + u, w = std::min(nb, p - kb);
                cpx<R>* L = num.factor + F.Loff;
                cpx<R>* U = num.factor + F.Uoff;
                const cpx<R>* Li = num.tiles + F.tileoff + (kb / nb) * 2 * nb * nb;
                const cpx<R>* Ui = Li + nb * nb;
                const int32_t* pr = num.perm + F.pivoff + kb;
                // row panels: rows kb..kb+w of the remaining pivot columns and of U12
                num.fplan.add(Li, nb, L + (kb + w) * ld + kb, ld, L + (kb + w) * ld + kb, ld, w, p - kb - w, w,
                              WT_GEMM_SET, pr);
                num.fplan.add(Li, nb, U + kb, p, U + kb, p, w, u, w, WT_GEMM_SET, pr);
                // column panel below the tile (L11 and L21 rows)
                num.fplan.add(L + kb * ld + kb + w, ld, Ui, nb, L + kb * ld + kb + w, ld, p + u - kb - w, w, w,
                              WT_GEMM_SET);
            }
            int g = num.fplan.end();
            if (g >= 0) num.fsteps.push_back({WT_STEP_GEMM, 0, 0, 0, par, g});
            // updates by this tile: inside the current super-panel of NB2
            // pivots (all rows), and beyond it for the super-panel's own rows
            // (remaining pivot columns and U12)
            num.fplan.begin();
            for (int i = 0; i < cnt; ++i) {
                const wt_front& F = S.fronts[lst[i]];
                const int64_t p = F.p, u = F.u, ld = p + u, w = std::min(nb, p - kb);
                const int64_t K1 = std::min((kb / NB2 + 1) * NB2, p);
                cpx<R>* L = num.factor + F.Loff;
                cpx<R>* U = num.factor + F.Uoff;
                num.fplan.add(L + kb * ld + kb + w, ld, L + (kb + w) * ld + kb, ld, L + (kb + w) * ld + kb + w, ld,
                              p + u - kb - w, K1 - kb - w, w, WT_GEMM_SUB);
                num.fplan.add(L + kb * ld + kb + w, ld, L + K1 * ld + kb, ld, L + K1 * ld + kb + w, ld,
                              K1 - kb - w, p - K1, w, WT_GEMM_SUB);
                num.fplan.add(L + kb * ld + kb + w, ld, U + kb, p, U + kb + w, p, K1 - kb - w, u, w, WT_GEMM_SUB);
            }
            g = num.fplan.end();
            if (g >= 0) num.fsteps.push_back({WT_STEP_GEMM, 0, 0, 0, par, g});
            // super-panel complete: one rank-(K1-K0) update of everything to
            // the right of and below it
            num.fplan.begin();
            for (int i = 0; i < cnt; ++i) {
                const wt_front& F = S.fronts[lst[i]];
                const int64_t p = F.p, u = F.u, ld = p + u, w = std::min(nb, p - kb);
                const int64_t K0 = (kb / NB2) * NB2, K1 = std::min(K0 + NB2, p);
                if (kb + w != K1 || K1 >= p) continue;
                cpx<R>* L = num.factor + F.Loff;
                cpx<R>* U = num.factor + F.Uoff;
                num.fplan.add(L + K0 * ld + K1, ld, L + K1 * ld + K0, ld, L + K1 * ld + K1, ld, p + u - K1, p - K1,
                              K1 - K0, WT_GEMM_SUB);
                num.fplan.add(L + K0 * ld + K1, ld, U + K0, p, U + K1, p, p - K1, u, K1 - K0, WT_GEMM_SUB);
            }
            g = num.fplan.end();
            if (g >= 0) num.fsteps.push_back({WT_STEP_GEMM, 0, 0, 0, par, g});
        }
        // contribution blocks: CB = -L21 U12, then the children's part
        num.fplan.begin();
        for (int f : S.levels[l]) {
            const wt_front& F = S.fronts[f];
            if (F.u == 0) continue;
            const int64_t p = F.p, u = F.u, ld = p + u;
            num.fplan.add(num.factor + F.Loff + p, ld, num.factor + F.Uoff, p, num.cb[par] + F.CBoff, u, u, u, p,
                          WT_GEMM_NSET);
        }
        int g = num.fplan.end();
        if (g >= 0) num.fsteps.push_back({WT_STEP_GEMM, 0, 0, 0, par, g});
        extend_add(l, 2);
    }
    return num.fplan.upload();
}

template <typename R>
static int alloc_numeric(wt_solver_impl& s, wt_numeric<R>& num)
{
    if (num.ready) return WT_OK;
    const wt_symbolic& S = s.S;
    int rc;
    if (num.ext) {
        if (!num.factor || !num.cb[0] || !num.cb[1] || !num.tiles || !num.perm || !num.status) return WT_ERR_STATE;
    } else if ((rc = dalloc(&num.factor, S.factor_len)) || (rc = dalloc(&num.cb[0], S.cb_len[0])) ||
               (rc = dalloc(&num.cb[1], S.cb_len[1])) || (rc = dalloc(&num.tiles, S.tile_len)) ||
               (rc = dalloc(&num.perm, S.piv_len)) || (rc = dalloc(&num.status, WT_PIVSTAT_LEN))) {
        num.release();
        return rc;
    }
    if ((rc = build_factor_plan(s, num))) {
        num.release();
        return rc;
    }
    num.ready = true;
    return WT_OK;
}

// Enqueues the whole numeric factorization on the solver stream.  When
// gemm_ms is given, every GEMM launch is bracketed by events (ev, two per
// launch index) for the autotuner.
template <typename R>
static int enqueue_factor(wt_solver_impl& s, wt_numeric<R>& num, std::vector<hipEvent_t>* ev)
{
    hipStream_t st = s.stream;
    const wt_symbolic& S = s.S;
    wt_prof& P = s.prof;
    const double cs = sizeof(cpx<R>);
    int rc;
    if ((rc = wt_pivstat_init(num.status, st))) return rc;
    if ((rc = P.run(st, WT_PROF_FILL, 0, S.factor_len * cs,
                    [&] { return wt_fill_zero<R>(num.factor, S.factor_len, st); })))
        return rc;
    if ((rc = P.run(st, WT_PROF_FILL, 0, s.nnz * 16.0 + s.n * 12.0,
                    [&] { return wt_row_scale(s.rowptr, s.vals, s.n, s.rscale, st); })))
        return rc;
    if ((rc = P.run(st, WT_PROF_FILL, 0, s.nnz * (16.0 + 8.0 + cs),
                    [&] { return wt_scatter_rows<R>(s.rowptr, s.vals, s.dst, s.rscale, s.n, num.factor, st); })))
        return rc;
    for (const wt_step& t : num.fsteps) {
        switch (t.kind) {
        case WT_STEP_ZERO:
            rc = P.run(st, WT_PROF_FILL, 0, t.a * cs, [&] { return wt_fill_zero<R>(num.cb[t.buf], t.a, st); });
            break;
        case WT_STEP_EXTADD:
            rc = P.run(st, WT_PROF_EXTADD, 0, 0, [&] {
                return wt_extend_add<R>(s.fronts, s.lists + t.list_off, s.cols + t.list_off, t.count, t.a, s.relmap,
                                        num.cb[t.buf ^ 1], num.factor, num.cb[t.buf], t.part, st);
            });
            break;
        case WT_STEP_DIAG:
            rc = P.run(st, WT_PROF_DIAG, 0, 0, [&] {
                return wt_diag_tiles<R>(s.fronts, s.lists + t.list_off, t.count, t.a, num.factor, num.tiles,
                                        num.perm, num.status, st);
            });
            break;
        case WT_STEP_GEMM:
            if (ev) (void)hipEventRecord((*ev)[2 * t.gemm], st);
        {
            const int v = num.fplan.launches[t.gemm].var;
            const int cls = v == 0 ? WT_PROF_GEMM_FACTOR
                                   : (v == WT_GEMM_SPLIT && sizeof(R) == 4 ? WT_PROF_GEMM_FACTOR_SPLIT
                                                                           : WT_PROF_GEMM_FACTOR_NARROW);
            rc = P.run(st, cls, num.fplan.launches[t.gemm].flops, 0, [&] { return num.fplan.run(t.gemm, st); });
        }
            if (ev) (void)hipEventRecord((*ev)[2 * t.gemm + 1], st);
            break;
        default: rc = WT_ERR_STATE;
        }
        if (rc) return rc;
    }
    return WT_OK;
}

template <typename R>
static int factor_prepare(wt_solver_impl& s, wt_numeric<R>& num)
{
    if (!s.have_vals) return WT_ERR_STATE;
    int rc = alloc_numeric(s, num);
    if (rc) return rc;
    if (!s.rscale && (rc = dalloc(&s.rscale, s.n))) return rc;
    return WT_OK;
}

template <typename R>
static int factor_finish(wt_solver_impl& s, wt_numeric<R>& num)
{
    num.have_pivstat = false;
    WT_HIP(hipMemcpyAsync(num.pivstat, num.status, sizeof(num.pivstat), hipMemcpyDeviceToHost, s.stream));
    WT_HIP(hipStreamSynchronize(s.stream));
    num.have_pivstat = true;
    s.prof.flush();
    s.factored_prec = sizeof(R) == 4 ? 32 : 64;
    return num.pivstat[0] ? WT_ERR_PIVOT : WT_OK;
}

template <typename R>
int wt_solver_factor_launch(wt_solver_impl& s, wt_numeric<R>& num);
template <typename R>
static int tune_solve_plan(wt_solver_impl& s, wt_numeric<R>& num, int nrhs);

// With s.use_graphs (and profiling off) the sequence is captured once into
// a HIP graph and replayed: thousands of small launches per factorization
// of a batch of small models become one graph launch.
template <typename R>
int wt_solver_factor_run(wt_solver_impl& s, wt_numeric<R>& num)
{
    int rc = wt_solver_factor_launch(s, num);
    return rc ? rc : factor_finish(s, num);
}

template <typename R>
int wt_solver_factor_wait(wt_solver_impl& s, wt_numeric<R>& num)
{
    return factor_finish(s, num);
}

// Enqueues the factorization (graph or eager) without waiting for it.
template <typename R>
int wt_solver_factor_launch(wt_solver_impl& s, wt_numeric<R>& num)
{
    int rc = factor_prepare(s, num);
    if (rc) return rc;
    if (s.use_graphs && !s.prof.on) {
        if (!num.fgraph) {
            wt_rt_capture excl;
            WT_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_factor(s, num, nullptr);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s.stream, &g);
            if (rc == WT_OK && e != hipSuccess) rc = WT_ERR_HIP;
            if (rc == WT_OK && hipGraphInstantiate(&num.fgraph, g, nullptr, nullptr, 0) != hipSuccess) {
                num.fgraph = nullptr;
                rc = WT_ERR_HIP;
            }
            if (g) (void)hipGraphDestroy(g);
            if (rc) return rc;
        }
        WT_HIP(hipGraphLaunch(num.fgraph, s.stream));
        return WT_OK;
    }
    return enqueue_factor(s, num, nullptr);
}

// Tile-shape autotuner: one full factorization per GEMM tile variant, each
// launch timed with events; every launch then keeps its fastest variant.
// Each run is a complete factorization (the variants compute identical
// results), so the factor is valid afterwards.  *speedup: summed GEMM time
// under the default choice over the summed time under the tuned choice.
template <typename R>
int wt_solver_autotune(wt_solver_impl& s, wt_numeric<R>& num, double* speedup)
{
    int rc = factor_prepare(s, num);
    if (rc) return rc;
    const size_t nl = num.fplan.launches.size();
    std::vector<int> dflt(nl), best(nl);
    std::vector<float> tbest(nl, INFINITY), tdef(nl, 0.f);
    for (size_t i = 0; i < nl; ++i) dflt[i] = best[i] = num.fplan.launches[i].var;
    std::vector<hipEvent_t> ev(2 * nl, nullptr);
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = WT_ERR_HIP;
    const bool prof = s.prof.on;
    s.prof.on = false;
    for (int v = 0; v < (s.gemm_split ? wt_gemm_nvar<R>() : WT_GEMM_SPLIT) && rc == WT_OK; ++v) {
        for (auto& L : num.fplan.launches) L.var = v;
        if ((rc = enqueue_factor(s, num, &ev)) || (rc = factor_finish(s, num))) break;
        for (size_t i = 0; i < nl; ++i) {
            float t = 0;
            if (hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]) != hipSuccess) continue;
            if (v == dflt[i]) tdef[i] = t;
            if (t < tbest[i]) { tbest[i] = t; best[i] = v; }
        }
    }
    s.prof.on = prof;
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    for (size_t i = 0; i < nl; ++i) num.fplan.launches[i].var = rc == WT_OK ? best[i] : dflt[i];
    num.drop_graphs();
    num.tuned = rc == WT_OK;
    // a solve plan built earlier is tuned now; later ones when they are built
    if (rc == WT_OK && num.snrhs > 0) rc = tune_solve_plan(s, num, num.snrhs);
    if (rc == WT_OK && speedup) {
        double a = 0, b = 0;
        for (size_t i = 0; i < nl; ++i) { a += tdef[i]; b += std::isfinite(tbest[i]) ? tbest[i] : tdef[i]; }
        *speedup = b > 0 ? a / b : 1.0;
    }
    return rc;
}

template <typename R>
static int build_solve_plan(wt_solver_impl& s, wt_numeric<R>& num, int nrhs)
{
    const wt_symbolic& S = s.S;
    const int nl = (int)S.levels.size();
    const int64_t n = S.n, NB2 = WT_NB2;
    if (num.sgraph) (void)hipGraphExecDestroy(num.sgraph);
    num.sgraph = nullptr;
    num.sgraph_nrhs = 0;
    if (!num.ext) {
        dfree(num.y);
        dfree(num.w[0]);
        dfree(num.w[1]);
    }
    dfree(num.d_sp);
    num.hsp.clear();
    num.splan.release();
    num.splan.allow_split = s.gemm_split;
    num.fwd.clear();
    num.bwd.clear();
    int rc;
    if (num.ext) {
        if (!num.y || !num.w[0] || !num.w[1]) return WT_ERR_STATE;
    } else if ((rc = dalloc(&num.y, n * nrhs)) || (rc = dalloc(&num.w[0], S.w_len[0] * nrhs)) ||
               (rc = dalloc(&num.w[1], S.w_len[1] * nrhs))) {
        return rc;
    }
    cpx<R>* y = num.y;
    auto nsp = [&](const wt_front& F) { return (F.p + NB2 - 1) / NB2; };
    // forward substitution, bottom-up; per super-panel: the panel's tiles in
    // one launch, then the rows below it and the update rows W by GEMM
    for (int l = nl - 1; l >= 0; --l) {
        const int par = l & 1;
        int64_t wl = 0;
        for (int f : S.levels[l]) wl = std::max(wl, S.fronts[f].Woff + S.fronts[f].u);
        if (wl > 0) num.fwd.push_back({WT_STEP_ZERO, 0, wl * nrhs, 0, par, -1});
        if (l + 1 < nl)
            for (int k = 0; k < 2; ++k)
                if (s.lev_kids_cnt[k][l + 1] > 0)
                    num.fwd.push_back({WT_STEP_SGATHER, s.lev_kids_cnt[k][l + 1], 0, s.lev_kids_off[k][l + 1], par, -1});
        const std::vector<int64_t>& ps = s.lev_sorted_p[l];
        const int32_t* lst = s.hlists.data() + s.lev_sorted_off[l];
        for (int64_t K0 = 0; K0 < ps[0]; K0 += NB2) {
            const int cnt = count_gt(ps, K0);
            const int64_t off = (int64_t)num.hsp.size();
            for (int i = 0; i < cnt; ++i) {
                const int64_t K1 = std::min(K0 + NB2, S.fronts[lst[i]].p);
                num.hsp.push_back(lst[i]);
                num.hsp.push_back((int32_t)K0);
                num.hsp.push_back((int32_t)K1);
            }
            num.fwd.push_back({WT_STEP_SPFWD, cnt, 0, off, par, -1});
            num.splan.begin();
            for (int i = 0; i < cnt; ++i) {
                const wt_front& F = S.fronts[lst[i]];
                const int64_t p = F.p, u = F.u, ld = p + u, K1 = std::min(K0 + NB2, p);
                const cpx<R>* L = num.factor + F.Loff;
                cpx<R>* ys = y + F.pstart + K0;
                num.splan.add(L + K0 * ld + K1, ld, ys, n, y + F.pstart + K1, n, p - K1, nrhs, K1 - K0, WT_GEMM_SUB);
                num.splan.add(L + K0 * ld + p, ld, ys, n, num.w[par] + F.Woff * nrhs, u, u, nrhs, K1 - K0,
                              WT_GEMM_SUB);
            }
            int g = num.splan.end();
            if (g >= 0) num.fwd.push_back({WT_STEP_GEMM, 0, 0, 0, par, g});
        }
    }
    // back substitution, top-down; super-panels from the last one back
    for (int l = 0; l < nl; ++l) {
        const int par = l & 1;
        if (s.lev_upd_cnt[l] > 0) {
            num.bwd.push_back({WT_STEP_SGATHER_UPD, s.lev_upd_cnt[l], 0, s.lev_upd_off[l], par, -1});
            // y_p -= U12 x_u: few output tiles (p rows, nrhs columns) and a
            // long K (u) at the top of the tree, so each product is split
            // along K into pieces of at least splitk_depth until the launch
            // has about WT_SPLITK_BLOCKS row tiles.  The rule depends on the
            // tree only (not on nrhs): every column is solved the same in
            // any block.
            int64_t mt = 0;
            for (int f : S.levels[l])
                if (S.fronts[f].u > 0) mt += (S.fronts[f].p + 63) / 64;
            num.splan.begin();
            for (int f : S.levels[l]) {
                const wt_front& F = S.fronts[f];
                if (F.u == 0) continue;
                int sk = 1;
                if (s.splitk_depth > 0 && mt > 0) {
                    const int64_t want = (WT_SPLITK_BLOCKS + mt - 1) / mt, cap = F.u / s.splitk_depth;
                    sk = (int)std::max<int64_t>(1, std::min<int64_t>(want, cap));
                }
                num.splan.add_split(num.factor + F.Uoff, F.p, num.w[par] + F.Woff * nrhs, F.u, y + F.pstart, n, F.p,
                                    nrhs, F.u, WT_GEMM_SUB, sk);
            }
            int g = num.splan.end();
            if (g >= 0) num.bwd.push_back({WT_STEP_GEMM, 0, 0, 0, par, g});
        }
        const std::vector<int64_t>& ps = s.lev_sorted_p[l];
        const int32_t* lst = s.hlists.data() + s.lev_sorted_off[l];
        const int64_t nsteps = (ps[0] + NB2 - 1) / NB2;
        for (int64_t st = 0; st < nsteps; ++st) {
            const int64_t off = (int64_t)num.hsp.size();
            int cnt = 0;
            for (int i = 0; i < (int)ps.size(); ++i) {
                const wt_front& F = S.fronts[lst[i]];
                const int64_t sp = nsp(F) - 1 - st;
                if (sp < 0) continue;
                num.hsp.push_back(lst[i]);
                num.hsp.push_back((int32_t)(sp * NB2));
                num.hsp.push_back((int32_t)std::min((sp + 1) * NB2, F.p));
                ++cnt;
            }
            if (cnt > 0) num.bwd.push_back({WT_STEP_SPBWD, cnt, 0, off, par, -1});
            num.splan.begin();
            for (int i = 0; i < (int)ps.size(); ++i) {
                const wt_front& F = S.fronts[lst[i]];
                const int64_t sp = nsp(F) - 1 - st;
                if (sp <= 0) continue;
                const int64_t p = F.p, ld = p + F.u, K0 = sp * NB2, K1 = std::min(K0 + NB2, p);
                const cpx<R>* L = num.factor + F.Loff;
                num.splan.add(L + K0 * ld, ld, y + F.pstart + K0, n, y + F.pstart, n, K0, nrhs, K1 - K0, WT_GEMM_SUB);
            }
            int g = num.splan.end();
            if (g >= 0) num.bwd.push_back({WT_STEP_GEMM, 0, 0, 0, par, g});
        }
    }
    if ((rc = dalloc(&num.d_sp, (int64_t)num.hsp.size()))) return rc;
    if (!num.hsp.empty())
        WT_HIP(rt_memcpy(num.d_sp, num.hsp.data(), num.hsp.size() * 4, hipMemcpyHostToDevice));
    if ((rc = num.splan.upload())) return rc;
    num.snrhs = nrhs;
    return WT_OK;
}

// Tile-shape tuning of the block solve's GEMM launches, as for the
// factorization: one solve per tile variant with every launch timed.  It
// runs on whatever y holds; the caller's next solve overwrites y first.
template <typename R>
static int tune_solve_plan(wt_solver_impl& s, wt_numeric<R>& num, int nrhs)
{
    const size_t nl = num.splan.launches.size();
    std::vector<int> best(nl);
    std::vector<float> tbest(nl, INFINITY);
    for (size_t i = 0; i < nl; ++i) best[i] = num.splan.launches[i].var;
    std::vector<hipEvent_t> ev(2 * nl, nullptr);
    int rc = WT_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = WT_ERR_HIP;
    const bool prof = s.prof.on;
    s.prof.on = false;
    for (int v = 0; v < (s.gemm_split ? wt_gemm_nvar<R>() : WT_GEMM_SPLIT) && rc == WT_OK; ++v) {
        for (auto& L : num.splan.launches) L.var = v;
        if ((rc = enqueue_solve(s, num, nrhs, &ev))) break;
        if (hipStreamSynchronize(s.stream) != hipSuccess) { rc = WT_ERR_HIP; break; }
        for (size_t i = 0; i < nl; ++i) {
            float t = 0;
            if (hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]) == hipSuccess && t < tbest[i]) {
                tbest[i] = t;
                best[i] = v;
            }
        }
    }
    s.prof.on = prof;
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    for (size_t i = 0; i < nl; ++i) num.splan.launches[i].var = best[i];
    if (num.sgraph) (void)hipGraphExecDestroy(num.sgraph);
    num.sgraph = nullptr;
    num.sgraph_nrhs = 0;
    return rc;
}

template <typename R>
int wt_solver_prepare_solve(wt_solver_impl& s, wt_numeric<R>& num, int nrhs)
{
    if (!num.ready) return WT_ERR_STATE;
    if (num.snrhs == nrhs) return WT_OK;
    int rc = build_solve_plan(s, num, nrhs);
    if (rc == WT_OK && num.tuned) rc = tune_solve_plan(s, num, nrhs);
    return rc;
}

// In-place block solve of A x = y for nrhs columns in elimination order
// (y column-major, ld = n).  Uses the solver's own y buffer when y == null.
template <typename R>
static int enqueue_solve(wt_solver_impl& s, wt_numeric<R>& num, int nrhs, std::vector<hipEvent_t>* ev);

template <typename R>
int wt_solver_solve_block(wt_solver_impl& s, wt_numeric<R>& num, int nrhs, cpx<R>* /*unused*/)
{
    if (!num.ready) return WT_ERR_STATE;
    int rc;
    if (num.snrhs != nrhs && (rc = wt_solver_prepare_solve(s, num, nrhs))) return rc;
    return enqueue_solve(s, num, nrhs, nullptr);
}

// The block solve's launches; with ev, every GEMM launch is bracketed by
// events (two per launch index) for the autotuner.
template <typename R>
static int enqueue_solve(wt_solver_impl& s, wt_numeric<R>& num, int nrhs, std::vector<hipEvent_t>* ev)
{
    int rc;
    hipStream_t st = s.stream;
    wt_prof& P = s.prof;
    for (int pass = 0; pass < 2; ++pass) {
        for (const wt_step& t : pass ? num.bwd : num.fwd) {
            switch (t.kind) {
            case WT_STEP_ZERO:
                rc = P.run(st, WT_PROF_FILL, 0, t.a * (double)sizeof(cpx<R>),
                           [&] { return wt_fill_zero<R>(num.w[t.buf], t.a, st); });
                break;
            case WT_STEP_SGATHER:
                rc = P.run(st, WT_PROF_PERMUTE, 0, 0, [&] {
                    return wt_solve_gather<R>(s.fronts, s.lists + t.list_off, t.count, s.relmap, num.w[t.buf ^ 1],
                                              num.w[t.buf], num.y, s.S.n, nrhs, st);
                });
                break;
            case WT_STEP_SGATHER_UPD:
                rc = P.run(st, WT_PROF_PERMUTE, 0, 0, [&] {
                    return wt_solve_gather_upd<R>(s.fronts, s.lists + t.list_off, t.count, s.upd, num.y,
                                                  num.w[t.buf], s.S.n, nrhs, st);
                });
                break;
            case WT_STEP_GEMM:
                if (ev) (void)hipEventRecord((*ev)[2 * t.gemm], st);
                rc = P.run(st, WT_PROF_GEMM_SOLVE, num.splan.launches[t.gemm].flops, 0,
                           [&] { return num.splan.run(t.gemm, st); });
                if (ev) (void)hipEventRecord((*ev)[2 * t.gemm + 1], st);
                break;
            case WT_STEP_SPFWD:
            case WT_STEP_SPBWD:
                rc = P.run(st, WT_PROF_TRSM, 0, 0, [&] {
                    return wt_sp_solve<R>(t.kind == WT_STEP_SPFWD, s.fronts, num.d_sp + t.list_off, t.count, num.factor,
                                          num.tiles, num.perm, num.y, s.S.n, nrhs, st);
                });
                break;
            default: rc = WT_ERR_STATE;
            }
            if (rc) return rc;
        }
    }
    return WT_OK;
}

// Tile variant of every GEMM launch of the factorization (which = 0) or of
// the block solve for nrhs columns (which = 1).  get: vars may be null to
// query the count; set: the count must match the plan.
template <typename R>
int wt_solver_plan_tiles(wt_solver_impl& s, wt_numeric<R>& num, int which, int nrhs, int32_t* vars,
                         const int32_t* in, int64_t* count)
{
    int rc;
    if (which == 0) {
        if ((rc = factor_prepare(s, num))) return rc;
    } else if (which == 1) {
        if (!num.ready) return WT_ERR_STATE;
        if (nrhs < 1 || nrhs > 64) return WT_ERR_ARG;
        if ((rc = wt_solver_prepare_solve(s, num, nrhs))) return rc;
    } else {
        return WT_ERR_ARG;
    }
    auto& L = which == 0 ? num.fplan.launches : num.splan.launches;
    if (in) {
        if (*count != (int64_t)L.size()) return WT_ERR_ARG;
        for (int64_t i = 0; i < *count; ++i)
            if (in[i] < 0 || in[i] >= WT_GEMM_NVAR) return WT_ERR_ARG;
        for (int64_t i = 0; i < *count; ++i) L[i].var = in[i];
        if (which == 0) {
            if (num.fgraph) (void)hipGraphExecDestroy(num.fgraph);
            num.fgraph = nullptr;
        } else {
            if (num.sgraph) (void)hipGraphExecDestroy(num.sgraph);
            num.sgraph = nullptr;
            num.sgraph_nrhs = 0;
        }
        return WT_OK;
    }
    *count = (int64_t)L.size();
    if (vars)
        for (size_t i = 0; i < L.size(); ++i) vars[i] = L[i].var;
    return WT_OK;
}

template int wt_solver_plan_tiles<float>(wt_solver_impl&, wt_numeric<float>&, int, int, int32_t*, const int32_t*,
                                         int64_t*);
template int wt_solver_plan_tiles<double>(wt_solver_impl&, wt_numeric<double>&, int, int, int32_t*, const int32_t*,
                                          int64_t*);
template void wt_numeric<float>::drop_graphs();
template void wt_numeric<double>::drop_graphs();
template int wt_solver_factor_run<float>(wt_solver_impl&, wt_numeric<float>&);
template int wt_solver_factor_run<double>(wt_solver_impl&, wt_numeric<double>&);
template int wt_solver_factor_launch<float>(wt_solver_impl&, wt_numeric<float>&);
template int wt_solver_factor_launch<double>(wt_solver_impl&, wt_numeric<double>&);
template int wt_solver_factor_wait<float>(wt_solver_impl&, wt_numeric<float>&);
template int wt_solver_factor_wait<double>(wt_solver_impl&, wt_numeric<double>&);
template int wt_solver_autotune<float>(wt_solver_impl&, wt_numeric<float>&, double*);
template int wt_solver_autotune<double>(wt_solver_impl&, wt_numeric<double>&, double*);
template int wt_solver_prepare_solve<float>(wt_solver_impl&, wt_numeric<float>&, int);
template int wt_solver_prepare_solve<double>(wt_solver_impl&, wt_numeric<double>&, int);
template int wt_solver_solve_block<float>(wt_solver_impl&, wt_numeric<float>&, int, cpx<float>*);
template int wt_solver_solve_block<double>(wt_solver_impl&, wt_numeric<double>&, int, cpx<double>*);
