// C ABI of the batched solver: many independent small models, each with its
// own grid, medium and frequency, assembled, factored and solved together.
//
// The reference solves one model at one frequency per pass of its driver
// loop (elwave.f90:147-186: DISPERSION, ASMBLE, STRUMPACK_SERIAL_FINTER job
// 2, then job 3 per source).  Here the models of a batch form one
// block-diagonal system whose elimination forest advances level by level,
// so every launch of the factorization and of the solve carries the fronts
// of all models at that depth.  Models are cut into groups whose CSR stays
// below 2^31 entries (32-bit row pointers); each group is one such system
// with its own stream, and the groups' factorizations run concurrently.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "wt_rt.h"
#include "wt_solver.h"
#include "wt_vec.h"

int64_t wt_nnz(int nx, int ny, int nz);
template <typename R>
int wt_launch_values(const wt_grid_params& p, const void* d_lam, const void* d_mu, const void* d_rho, void* d_vals,
                     int* d_err, hipStream_t stream);

namespace {

constexpr int kSegBlocks = 8;  // blocks per model in the segmented norms

struct group {
    wt_solver_impl s;
    int m0 = 0, m1 = 0;             // models [m0, m1)
    std::vector<int64_t> unk_off;   // local unknown offset per model, + end
    std::vector<int64_t> node_off;  // local node offset per model, + end
    std::vector<int64_t> nnz_off;   // local CSR offset per model, + end
    int64_t* d_seg = nullptr;       // unk_off on the device
    int32_t* d_ncol = nullptr;      // right-hand sides per model (device)
    double* seg_part = nullptr;
    double* seg_out = nullptr;
    int seg_cap = 0;
    int nm() const { return m1 - m0; }
};

template <typename T>
int dalloc(T** p, int64_t count)
{
    *p = nullptr;
    if (rt_malloc((void**)p, (size_t)(count > 0 ? count : 1) * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return WT_ERR_NOMEM;
    }
    return WT_OK;
}

void free_group(group* g)
{
    if (!g) return;
    (void)rt_free(g->d_seg);
    (void)rt_free(g->d_ncol);
    (void)rt_free(g->seg_part);
    (void)rt_free(g->seg_out);
    wt_solver_free(g->s);
    delete g;
}

// Device bytes of the numeric state of a symbolic analysis: factor,
// contribution blocks, inverse tiles, pivots, solve and refinement blocks.
int64_t numeric_bytes(const wt_symbolic& S, int64_t cs, int64_t nrhs)
{
    return (S.factor_len + S.cb_len[0] + S.cb_len[1] + S.tile_len) * cs + S.piv_len * 4 +
           (S.n + S.w_len[0] + S.w_len[1]) * nrhs * cs + 3 * S.n * nrhs * 16;
}

int seg_buffers(group& g, int nrhs)
{
    if (g.seg_cap >= nrhs) return WT_OK;
    (void)rt_free(g.seg_part);
    (void)rt_free(g.seg_out);
    g.seg_part = g.seg_out = nullptr;
    g.seg_cap = 0;
    int rc;
    if ((rc = dalloc(&g.seg_part, (int64_t)g.nm() * kSegBlocks * nrhs)) ||
        (rc = dalloc(&g.seg_out, (int64_t)g.nm() * nrhs)))
        return rc;
    g.seg_cap = nrhs;
    return WT_OK;
}

// One refinement step on the group's buffers: d = A_f^{-1} r, x += d,
// r = b - A x (fp64), per-model column norms of r.
template <typename R>
int refine_step(group& g, wt_numeric<R>& num, int nrhs)
{
    wt_solver_impl& s = g.s;
    hipStream_t q = s.stream;
    const int64_t n = s.n;
    int rc;
    if ((rc = wt_to_elim<double, R>(s.rb, s.iperm, s.rscale, num.y, n, nrhs, q))) return rc;
    if ((rc = wt_solver_solve_block<R>(s, num, nrhs, nullptr))) return rc;
    if ((rc = wt_from_elim<R, double>(num.y, s.iperm, s.xb, n, nrhs, 1, q))) return rc;
    if ((rc = wt_residual<double>(s.rowptr, s.colidx, s.vals, s.xb, s.bb, s.rb, s.partial, s.norms, n, nrhs, q)))
        return rc;
    return wt_seg_colnorm2(s.rb, g.d_seg, g.nm(), nrhs, g.seg_part, kSegBlocks, g.seg_out, q);
}

// Refinement of the group's block s.bb (row-major n x nrhs, columns beyond
// each model's count already zero) into s.xb until every model's columns
// satisfy |r_j| / |b_j| <= tol, or max_iter steps.
template <typename R>
int refine_group(group& g, wt_numeric<R>& num, int nrhs, const int32_t* ncol, double tol, int max_iter,
                 wt_solve_stats* st)
{
    wt_solver_impl& s = g.s;
    hipStream_t q = s.stream;
    const int64_t n = s.n;
    const int nm = g.nm();
    int rc;
    if ((rc = seg_buffers(g, nrhs))) return rc;
    if ((rc = wt_solver_prepare_solve<R>(s, num, nrhs))) return rc;
    std::vector<double> hb((size_t)nm * nrhs), hr((size_t)nm * nrhs);
    if ((rc = wt_seg_colnorm2(s.bb, g.d_seg, nm, nrhs, g.seg_part, kSegBlocks, g.seg_out, q))) return rc;
    WT_HIP(hipMemcpyAsync(hb.data(), g.seg_out, hb.size() * 8, hipMemcpyDeviceToHost, q));
    WT_HIP(hipMemsetAsync(s.xb, 0, n * nrhs * sizeof(cpx<double>), q));
    if ((rc = wt_copy<double>(s.bb, s.rb, n * nrhs, q))) return rc;
    const bool graph = s.use_graphs && !s.prof.on;
    if (graph && (!num.sgraph || num.sgraph_nrhs != nrhs)) {
        if (num.sgraph) (void)hipGraphExecDestroy(num.sgraph);
        num.sgraph = nullptr;
        wt_rt_capture excl;
        WT_HIP(hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal));
        rc = refine_step<R>(g, num, nrhs);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(q, &gr);
        if (rc == WT_OK && e != hipSuccess) rc = WT_ERR_HIP;
        if (rc == WT_OK && hipGraphInstantiate(&num.sgraph, gr, nullptr, nullptr, 0) != hipSuccess) {
            num.sgraph = nullptr;
            rc = WT_ERR_HIP;
        }
        if (gr) (void)hipGraphDestroy(gr);
        if (rc) return rc;
        num.sgraph_nrhs = nrhs;
    }
    std::vector<double> first(nm, INFINITY), last(nm, INFINITY), absr(nm, 0.0), prev(nm, INFINITY);
    std::vector<char> stall(nm, 0);
    int it = 0;
    bool bad = false;
    while (it < max_iter) {
        if (graph) WT_HIP(hipGraphLaunch(num.sgraph, q));
        else if ((rc = refine_step<R>(g, num, nrhs))) return rc;
        WT_HIP(hipMemcpyAsync(hr.data(), g.seg_out, hr.size() * 8, hipMemcpyDeviceToHost, q));
        WT_HIP(hipStreamSynchronize(q));
        s.prof.flush();
        ++it;
        bool done = true, moving = false;
        for (int k = 0; k < nm; ++k) {
            double rel = 0.0, am = 0.0;
            for (int c = 0; c < ncol[k]; ++c) {
                const double rr = std::sqrt(hr[(size_t)k * nrhs + c]), bn = std::sqrt(hb[(size_t)k * nrhs + c]);
                const double v = bn > 0 ? rr / bn : rr;
                if (!(v <= rel)) rel = v;  // NaN propagates
                if (rr > am) am = rr;
            }
            if (it == 1) first[k] = rel;
            last[k] = rel;
            absr[k] = am;
            if (!std::isfinite(rel)) bad = true;
            if (!(rel <= tol)) {
                done = false;
                // a model whose residual stopped falling is not refined on
                // its own account; the loop ends when no model is moving
                if (rel >= prev[k]) stall[k] = 1;
                else moving = true;
            }
            prev[k] = rel;
        }
        if (done || bad || (tol > 0 && !moving)) break;
    }
    bool conv_all = true;
    for (int k = 0; k < nm; ++k) {
        wt_solve_stats& o = st[k];
        o.iterations = it;
        o.nrhs = ncol[k];
        o.rel_residual = last[k];
        o.rel_residual_first = first[k];
        o.abs_residual = absr[k];
        o.ms = 0;
        o.converged = (tol <= 0 || last[k] <= tol) ? 1 : 0;
        o.stalled = (o.converged || !stall[k]) ? 0 : 1;
        if (!o.converged) conv_all = false;
    }
    if (bad) return WT_ERR_PIVOT;
    return conv_all ? WT_OK : WT_ERR_NOCONV;
}

}  // namespace

// Numeric buffers shared by the groups of wt_batch_run (one group resident
// at a time), grown to the largest group's needs and kept across runs.
struct num_pool {
    static constexpr int NB = 9;  // factor, cb0, cb1, tiles, perm, status, y, w0, w1
    void* p[NB] = {};
    int64_t cap[NB] = {};
    int ensure(const int64_t* need)
    {
        for (int i = 0; i < NB; ++i) {
            if (cap[i] >= need[i] && p[i]) continue;
            (void)rt_free(p[i]);
            p[i] = nullptr;
            cap[i] = 0;
            if (rt_malloc(&p[i], (size_t)(need[i] > 0 ? need[i] : 16)) != hipSuccess) {
                p[i] = nullptr;
                return WT_ERR_NOMEM;
            }
            cap[i] = need[i] > 0 ? need[i] : 16;
        }
        return WT_OK;
    }
    void release()
    {
        for (int i = 0; i < NB; ++i) {
            (void)rt_free(p[i]);
            p[i] = nullptr;
            cap[i] = 0;
        }
    }
};

struct wt_batch {
    int nmodels = 0;
    std::vector<wt_grid_dims> dims;
    std::vector<int64_t> unk_off, node_off;  // global, + end
    std::vector<group*> groups;
    std::vector<int> group_of;
    bool have_model = false;
    num_pool pool;
};

namespace {
template <typename R>
wt_numeric<R>& numeric_of(wt_solver_impl& s);
template <>
wt_numeric<float>& numeric_of<float>(wt_solver_impl& s) { return s.f32; }
template <>
wt_numeric<double>& numeric_of<double>(wt_solver_impl& s) { return s.f64; }

// Points a group's numeric state at the shared pool.
template <typename R>
void lend(num_pool& P, wt_numeric<R>& num)
{
    if (num.ready && !num.ext) num.release();
    num.ext = true;
    num.factor = (cpx<R>*)P.p[0];
    num.cb[0] = (cpx<R>*)P.p[1];
    num.cb[1] = (cpx<R>*)P.p[2];
    num.tiles = (cpx<R>*)P.p[3];
    num.perm = (int32_t*)P.p[4];
    num.status = (int*)P.p[5];
    num.y = (cpx<R>*)P.p[6];
    num.w[0] = (cpx<R>*)P.p[7];
    num.w[1] = (cpx<R>*)P.p[8];
}

template <typename R>
int run_groups(wt_batch* b, int nrhs, const int32_t* ncol, const void* rhs, void* x, int on_device, double tol,
               int max_iter, wt_solve_stats* stats)
{
    const int64_t cs = sizeof(cpx<R>);
    int64_t need[num_pool::NB] = {};
    for (group* g : b->groups) {
        const wt_symbolic& S = g->s.S;
        const int64_t v[num_pool::NB] = {S.factor_len * cs, S.cb_len[0] * cs, S.cb_len[1] * cs, S.tile_len * cs,
                                         S.piv_len * 4, WT_PIVSTAT_LEN * 4, S.n * nrhs * cs, S.w_len[0] * nrhs * cs,
                                         S.w_len[1] * nrhs * cs};
        for (int i = 0; i < num_pool::NB; ++i) need[i] = need[i] > v[i] ? need[i] : v[i];
    }
    int rc = b->pool.ensure(need);
    if (rc) return rc;
    const hipMemcpyKind in = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind outk = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    int result = WT_OK;
    for (group* g : b->groups) {
        wt_solver_impl& s = g->s;
        wt_numeric<R>& num = numeric_of<R>(s);
        if (!s.have_vals) return WT_ERR_STATE;
        lend(b->pool, num);
        rc = wt_solver_factor_run<R>(s, num);
        if (rc == WT_OK) rc = wt_solver_ir_buffers(s, nrhs);
        if (rc == WT_OK) {
            const int64_t row0 = b->unk_off[g->m0];
            WT_HIP(hipMemcpyAsync(s.bb, (const cpx<double>*)rhs + row0 * nrhs, s.n * nrhs * 16, in, s.stream));
            WT_HIP(hipMemcpyAsync(g->d_ncol, ncol + g->m0, g->nm() * 4, hipMemcpyHostToDevice, s.stream));
            rc = wt_seg_mask_cols(s.bb, g->d_seg, g->d_ncol, g->nm(), s.n, nrhs, s.stream);
            if (rc == WT_OK) rc = refine_group<R>(*g, num, nrhs, ncol + g->m0, tol, max_iter, stats + g->m0);
            if (rc == WT_OK || rc == WT_ERR_PIVOT || rc == WT_ERR_NOCONV) {
                WT_HIP(hipMemcpyAsync((cpx<double>*)x + row0 * nrhs, s.xb, s.n * nrhs * 16, outk, s.stream));
                WT_HIP(hipStreamSynchronize(s.stream));
            }
        }
        // the pool passes to the next group
        WT_HIP(hipStreamSynchronize(s.stream));
        num.release();
        if (rc == WT_ERR_PIVOT || (rc == WT_ERR_NOCONV && result != WT_ERR_PIVOT)) result = rc;
        else if (rc) return rc;
    }
    return result;
}
}  // namespace

extern "C" {

int wt_batch_destroy(wt_batch* b)
{
    if (!b) return WT_OK;
    for (group* g : b->groups) free_group(g);
    b->pool.release();
    delete b;
    return WT_OK;
}

int wt_batch_create(int nmodels, const int32_t* dims, int leaf_nodes, int64_t max_group_nnz, int64_t max_group_bytes,
                    wt_batch** out)
{
    if (!out || !dims || nmodels < 1) return WT_ERR_ARG;
    *out = nullptr;
    const int64_t cap = (1LL << 31) - 1;
    if (max_group_nnz <= 0 || max_group_nnz > cap) max_group_nnz = cap;
    if (leaf_nodes <= 0) leaf_nodes = 64;
    wt_batch* b = new (std::nothrow) wt_batch();
    if (!b) return WT_ERR_NOMEM;
    b->nmodels = nmodels;
    b->unk_off.assign(nmodels + 1, 0);
    b->node_off.assign(nmodels + 1, 0);
    b->group_of.assign(nmodels, 0);
    for (int k = 0; k < nmodels; ++k) {
        wt_grid_dims d{dims[3 * k], dims[3 * k + 1], dims[3 * k + 2]};
        if (d.nx < 2 || d.ny < 2 || d.nz < 2 || wt_nnz(d.nx, d.ny, d.nz) > max_group_nnz) {
            delete b;
            return WT_ERR_ARG;
        }
        b->dims.push_back(d);
        b->node_off[k + 1] = b->node_off[k] + (int64_t)d.nx * d.ny * d.nz;
        b->unk_off[k + 1] = 3 * b->node_off[k + 1];
    }
    // numeric-state bytes of each model (complex64 factor, 16 right-hand
    // sides) when groups are bounded by memory
    std::vector<int64_t> mb(nmodels, 0);
    if (max_group_bytes > 0)
        for (int k = 0; k < nmodels; ++k) {
            wt_symbolic S;
            int rc = wt_symbolic_build(b->dims[k].nx, b->dims[k].ny, b->dims[k].nz, leaf_nodes, WT_NB, nullptr,
                                       nullptr, S);
            if (rc) {
                delete b;
                return rc;
            }
            mb[k] = numeric_bytes(S, 8, 16);
        }
    // greedy groups of consecutive models
    int rc = WT_OK;
    for (int k = 0; k < nmodels && rc == WT_OK;) {
        group* g = new (std::nothrow) group();
        if (!g) { rc = WT_ERR_NOMEM; break; }
        b->groups.push_back(g);
        g->m0 = k;
        int64_t e = 0, by = 0;
        while (k < nmodels) {
            const int64_t ek = wt_nnz(b->dims[k].nx, b->dims[k].ny, b->dims[k].nz);
            if (k > g->m0 && (e + ek > max_group_nnz || (max_group_bytes > 0 && by + mb[k] > max_group_bytes)))
                break;
            e += ek;
            by += mb[k];
            b->group_of[k] = (int)b->groups.size() - 1;
            ++k;
        }
        g->m1 = k;
        std::vector<wt_grid_dims> gd(b->dims.begin() + g->m0, b->dims.begin() + g->m1);
        g->unk_off.assign(g->nm() + 1, 0);
        g->node_off.assign(g->nm() + 1, 0);
        g->nnz_off.assign(g->nm() + 1, 0);
        for (int i = 0; i < g->nm(); ++i) {
            const int64_t nodes = (int64_t)gd[i].nx * gd[i].ny * gd[i].nz;
            g->node_off[i + 1] = g->node_off[i] + nodes;
            g->unk_off[i + 1] = 3 * g->node_off[i + 1];
            g->nnz_off[i + 1] = g->nnz_off[i] + wt_nnz(gd[i].nx, gd[i].ny, gd[i].nz);
        }
        if ((rc = wt_solver_init_union(g->s, gd, leaf_nodes))) break;
        if ((rc = dalloc(&g->d_seg, g->nm() + 1)) || (rc = dalloc(&g->d_ncol, g->nm()))) break;
        if (rt_memcpy(g->d_seg, g->unk_off.data(), (g->nm() + 1) * 8, hipMemcpyHostToDevice) != hipSuccess)
            rc = WT_ERR_HIP;
        g->s.use_graphs = true;
    }
    if (rc) {
        wt_batch_destroy(b);
        return rc;
    }
    *out = b;
    return WT_OK;
}

int wt_batch_info(const wt_batch* b, int64_t* info, double* flops)
{
    if (!b || !info) return WT_ERR_ARG;
    int64_t fronts = 0, levels = 0, fl = 0, maxp = 0, cb = 0;
    double f = 0;
    for (const group* g : b->groups) {
        const wt_symbolic& S = g->s.S;
        fronts += (int64_t)S.fronts.size();
        levels = levels > (int64_t)S.levels.size() ? levels : (int64_t)S.levels.size();
        fl += S.factor_len;
        maxp = maxp > S.max_p ? maxp : S.max_p;
        cb += S.cb_len[0] + S.cb_len[1];
        f += S.flops;
    }
    int64_t nnz = 0;
    for (const group* g : b->groups) nnz += g->s.nnz;
    info[0] = b->unk_off[b->nmodels];
    info[1] = nnz;
    info[2] = fronts;
    info[3] = levels;
    info[4] = maxp;
    info[5] = (int64_t)b->groups.size();
    info[6] = fl;
    info[7] = cb;
    if (flops) *flops = f;
    return WT_OK;
}

int wt_batch_offsets(const wt_batch* b, int64_t* unk_off, int32_t* group_of)
{
    if (!b) return WT_ERR_ARG;
    if (unk_off) std::memcpy(unk_off, b->unk_off.data(), b->unk_off.size() * 8);
    if (group_of) for (int k = 0; k < b->nmodels; ++k) group_of[k] = b->group_of[k];
    return WT_OK;
}

int wt_batch_stream(const wt_batch* b, void** stream)
{
    if (!b || !stream || b->groups.empty()) return WT_ERR_ARG;
    *stream = (void*)b->groups[0]->s.stream;
    return WT_OK;
}

int wt_batch_set_graphs(wt_batch* b, int enable)
{
    if (!b) return WT_ERR_ARG;
    for (group* g : b->groups) g->s.use_graphs = enable != 0;
    return WT_OK;
}

int wt_batch_set_models(wt_batch* b, const void* lam, const void* mu, const double* rho, int on_device)
{
    if (!b || !lam || !mu || !rho) return WT_ERR_ARG;
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    for (group* g : b->groups) {
        wt_solver_impl& s = g->s;
        const int64_t nodes = s.n / 3, n0 = b->node_off[g->m0];
        int rc;
        if (!s.lam && ((rc = dalloc(&s.lam, nodes)) || (rc = dalloc(&s.mu, nodes)) || (rc = dalloc(&s.rho, nodes))))
            return rc;
        WT_HIP(hipMemcpyAsync(s.lam, (const cpx<double>*)lam + n0, nodes * 16, kind, s.stream));
        WT_HIP(hipMemcpyAsync(s.mu, (const cpx<double>*)mu + n0, nodes * 16, kind, s.stream));
        WT_HIP(hipMemcpyAsync(s.rho, rho + n0, nodes * 8, kind, s.stream));
    }
    for (group* g : b->groups) WT_HIP(hipStreamSynchronize(g->s.stream));
    b->have_model = true;
    return WT_OK;
}

int wt_batch_assemble(wt_batch* b, const wt_grid_params* params)
{
    if (!b || !params) return WT_ERR_ARG;
    if (!b->have_model) return WT_ERR_STATE;
    for (int k = 0; k < b->nmodels; ++k)
        if (params[k].nx != b->dims[k].nx || params[k].ny != b->dims[k].ny || params[k].nz != b->dims[k].nz)
            return WT_ERR_ARG;
    int rc;
    for (group* g : b->groups) {
        wt_solver_impl& s = g->s;
        if (!s.vals && (rc = dalloc(&s.vals, s.nnz))) return rc;
        if ((rc = wt_zero_i32(s.d_err, 1, s.stream))) return rc;
        for (int i = 0; i < g->nm(); ++i)
            if ((rc = wt_launch_values<double>(params[g->m0 + i], s.lam + g->node_off[i], s.mu + g->node_off[i],
                                               s.rho + g->node_off[i], s.vals + g->nnz_off[i], s.d_err, s.stream)))
                return rc;
    }
    int bad = 0;
    for (group* g : b->groups) {
        int he = 0;
        WT_HIP(hipMemcpyAsync(&he, g->s.d_err, sizeof(int), hipMemcpyDeviceToHost, g->s.stream));
        WT_HIP(hipStreamSynchronize(g->s.stream));
        bad |= he;
        g->s.have_vals = true;
        g->s.have_p = false;
    }
    return bad ? WT_ERR_ARG : WT_OK;
}

int wt_batch_factor(wt_batch* b, int prec)
{
    if (!b || (prec != WT_PREC_F32 && prec != WT_PREC_F64)) return WT_ERR_ARG;
    int rc = WT_OK, rw = WT_OK;
    for (group* g : b->groups) {
        rc = prec == WT_PREC_F32 ? wt_solver_factor_launch<float>(g->s, g->s.f32)
                                 : wt_solver_factor_launch<double>(g->s, g->s.f64);
        if (rc) break;
    }
    for (group* g : b->groups) {
        int r = prec == WT_PREC_F32 ? wt_solver_factor_wait<float>(g->s, g->s.f32)
                                    : wt_solver_factor_wait<double>(g->s, g->s.f64);
        if (r && rw == WT_OK) rw = r;
    }
    return rc ? rc : rw;
}

int wt_batch_autotune(wt_batch* b, int prec, double* speedup)
{
    if (!b || (prec != WT_PREC_F32 && prec != WT_PREC_F64)) return WT_ERR_ARG;
    double a = 0, w = 0;
    for (group* g : b->groups) {
        double sp = 1.0;
        int rc = prec == WT_PREC_F32 ? wt_solver_autotune<float>(g->s, g->s.f32, &sp)
                                     : wt_solver_autotune<double>(g->s, g->s.f64, &sp);
        if (rc) return rc;
        a += sp * (double)g->s.nnz;
        w += (double)g->s.nnz;
    }
    if (speedup) *speedup = w > 0 ? a / w : 1.0;
    return WT_OK;
}

int wt_batch_symbolic_info(int nmodels, const int32_t* dims, int leaf_nodes, int64_t* info, double* flops)
{
    if (!dims || !info || nmodels < 1) return WT_ERR_ARG;
    std::vector<wt_grid_dims> g(nmodels);
    for (int k = 0; k < nmodels; ++k) g[k] = wt_grid_dims{dims[3 * k], dims[3 * k + 1], dims[3 * k + 2]};
    wt_symbolic U;
    int rc = wt_union_build(g, leaf_nodes > 0 ? leaf_nodes : 64, U, nullptr, nullptr);
    if (rc) return rc;
    info[0] = U.n;
    info[1] = U.nnz;
    info[2] = (int64_t)U.fronts.size();
    info[3] = (int64_t)U.levels.size();
    info[4] = U.max_p;
    info[5] = U.factor_len;
    info[6] = U.cb_len[0] + U.cb_len[1];
    info[7] = U.w_len[0] + U.w_len[1];
    if (flops) *flops = U.flops;
    return WT_OK;
}

int wt_graph_host(int nx, int ny, int nz, int base, int32_t* rowptr, int32_t* colidx)
{
    return wt_pattern_host(nx, ny, nz, base, base, rowptr, colidx);
}

int wt_batch_run(wt_batch* b, int factor_prec, int nrhs, const int32_t* ncol, const void* rhs, void* x,
                 int on_device, double tol, int max_iter, wt_solve_stats* stats)
{
    if (!b || !ncol || !rhs || !x || !stats || nrhs < 1 || nrhs > 64 || max_iter < 1) return WT_ERR_ARG;
    if (factor_prec != WT_PREC_F32 && factor_prec != WT_PREC_F64) return WT_ERR_ARG;
    for (int k = 0; k < b->nmodels; ++k)
        if (ncol[k] < 0 || ncol[k] > nrhs) return WT_ERR_ARG;
    return factor_prec == WT_PREC_F32 ? run_groups<float>(b, nrhs, ncol, rhs, x, on_device, tol, max_iter, stats)
                                      : run_groups<double>(b, nrhs, ncol, rhs, x, on_device, tol, max_iter, stats);
}

int wt_batch_apply(wt_batch* b, int nrhs, const void* x, void* y, int on_device)
{
    if (!b || !x || !y || nrhs < 1 || nrhs > 64) return WT_ERR_ARG;
    const hipMemcpyKind in = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind outk = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    for (group* g : b->groups) {
        wt_solver_impl& s = g->s;
        if (!s.have_vals) return WT_ERR_STATE;
        int rc;
        if ((rc = wt_solver_ir_buffers(s, nrhs))) return rc;
        const int64_t row0 = b->unk_off[g->m0];
        WT_HIP(hipMemcpyAsync(s.xb, (const cpx<double>*)x + row0 * nrhs, s.n * nrhs * 16, in, s.stream));
        if ((rc = wt_residual<double>(s.rowptr, s.colidx, s.vals, s.xb, nullptr, s.rb, s.partial, s.norms, s.n, nrhs,
                                      s.stream)))
            return rc;
        WT_HIP(hipMemcpyAsync((cpx<double>*)y + row0 * nrhs, s.rb, s.n * nrhs * 16, outk, s.stream));
        WT_HIP(hipStreamSynchronize(s.stream));
    }
    return WT_OK;
}

int wt_batch_solve(wt_batch* b, int factor_prec, int nrhs, const int32_t* ncol, const void* rhs, void* x,
                   int on_device, double tol, int max_iter, wt_solve_stats* stats)
{
    if (!b || !ncol || !rhs || !x || !stats || nrhs < 1 || nrhs > 64 || max_iter < 1) return WT_ERR_ARG;
    const bool f32 = factor_prec == WT_PREC_F32;
    if (!f32 && factor_prec != WT_PREC_F64) return WT_ERR_ARG;
    for (int k = 0; k < b->nmodels; ++k)
        if (ncol[k] < 0 || ncol[k] > nrhs) return WT_ERR_ARG;
    const hipMemcpyKind in = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind outk = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    int result = WT_OK;
    for (group* g : b->groups) {
        wt_solver_impl& s = g->s;
        if (!s.have_vals) return WT_ERR_STATE;
        if ((f32 && !s.f32.ready) || (!f32 && !s.f64.ready)) return WT_ERR_STATE;
        int rc;
        if ((rc = wt_solver_ir_buffers(s, nrhs))) return rc;
        const int64_t row0 = b->unk_off[g->m0];
        WT_HIP(hipMemcpyAsync(s.bb, (const cpx<double>*)rhs + row0 * nrhs, s.n * nrhs * 16, in, s.stream));
        WT_HIP(hipMemcpyAsync(g->d_ncol, ncol + g->m0, g->nm() * 4, hipMemcpyHostToDevice, s.stream));
        if ((rc = wt_seg_mask_cols(s.bb, g->d_seg, g->d_ncol, g->nm(), s.n, nrhs, s.stream))) return rc;
        rc = f32 ? refine_group<float>(*g, s.f32, nrhs, ncol + g->m0, tol, max_iter, stats + g->m0)
                 : refine_group<double>(*g, s.f64, nrhs, ncol + g->m0, tol, max_iter, stats + g->m0);
        if (rc == WT_ERR_PIVOT || (rc == WT_ERR_NOCONV && result != WT_ERR_PIVOT)) result = rc;
        else if (rc) return rc;
        WT_HIP(hipMemcpyAsync((cpx<double>*)x + row0 * nrhs, s.xb, s.n * nrhs * 16, outk, s.stream));
        WT_HIP(hipStreamSynchronize(s.stream));
    }
    return result;
}

}  // extern "C"
