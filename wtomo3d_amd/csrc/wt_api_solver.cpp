// C ABI of the solver: handle API, iterative refinement, and the drop-in
// replacement of the reference's STRUMPACK shim (strumpack_finter.c:9-126).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "wt_solver.h"
#include "wt_rhs.h"
#include "wt_model.h"
#include "wt_rt.h"
#include "wt_vec.h"

int64_t wt_nnz(int nx, int ny, int nz);
template <typename R>
int wt_launch_values(const wt_grid_params& p, const void* d_lam, const void* d_mu, const void* d_rho, void* d_vals,
                     int* d_err, hipStream_t stream);

struct wt_solver {
    wt_solver_impl s;
};

namespace {

template <typename T>
int dalloc(T** p, int64_t count)
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
int salloc(T** p, int64_t count, hipStream_t st)
{
    return rt_malloc_async((void**)p, (size_t)(count > 0 ? count : 1) * sizeof(T), st) == hipSuccess ? WT_OK
                                                                                             : WT_ERR_NOMEM;
}

int ensure_ir(wt_solver_impl& s, int nrhs);
}  // namespace

int wt_solver_ir_buffers(wt_solver_impl& s, int nrhs) { return ensure_ir(s, nrhs); }

namespace {
int ensure_ir(wt_solver_impl& s, int nrhs)
{
    if (s.ir_cap >= nrhs) return WT_OK;
    (void)rt_free(s.xb); (void)rt_free(s.rb); (void)rt_free(s.bb); (void)rt_free(s.partial); (void)rt_free(s.norms);
    s.xb = s.rb = s.bb = nullptr;
    s.partial = s.norms = nullptr;
    s.ir_cap = 0;
    int64_t nb = wt_residual_blocks(s.n);
    int rc;
    if ((rc = dalloc(&s.xb, s.n * nrhs)) || (rc = dalloc(&s.rb, s.n * nrhs)) || (rc = dalloc(&s.bb, s.n * nrhs)) ||
        (rc = dalloc(&s.partial, nb * nrhs)) || (rc = dalloc(&s.norms, 2 * nrhs)))
        return rc;
    s.ir_cap = nrhs;
    return WT_OK;
}

// Refinement on a row-major block of <= 64 columns already on the device:
// x = 0, r = b; repeat { d = A_f^{-1} r (factor precision); x += d;
// r = b - A x (fp64) } until max_j |r_j| / |b_j| <= tol or max_iter solves,
// or a step that does not lower max_j |r_j| / |b_j| (tol > 0 only).
// Returns WT_ERR_NOCONV when it ends above tol > 0 (x and *st are written),
// WT_ERR_PIVOT for a non-finite residual.
template <typename R>
int refine(wt_solver_impl& s, wt_numeric<R>& num, int nrhs, const cpx<double>* b, cpx<double>* x, double tol,
           int max_iter, wt_solve_stats* st)
{
    hipStream_t q = s.stream;
    const int64_t n = s.n;
    int rc;
    const int64_t ncb = 1024;
    if ((rc = wt_colnorm2(b, n, nrhs, s.partial, ncb < wt_residual_blocks(n) ? ncb : wt_residual_blocks(n),
                          s.norms + nrhs, q)))
        return rc;
    if (max_iter < 1) WT_HIP(hipMemsetAsync(x, 0, n * nrhs * sizeof(cpx<double>), q));
    const cpx<double>* r = b;
    std::vector<double> h(2 * nrhs);
    int it = 0;
    double rel = INFINITY, first = INFINITY, prev = INFINITY;
    bool stalled = false;
    // For 17..64 columns the residual kernel writes the next solve's input
    // (scaled, in elimination order) itself, so r is never stored.
    const bool fused = wt_residual_elim_ok(n, nrhs);
    bool y_ready = false;
    while (it < max_iter) {
        wt_prof& P = s.prof;
        const double cs = sizeof(cpx<R>), blk = (double)n * nrhs;
        if ((rc = wt_solver_prepare_solve<R>(s, num, nrhs))) return rc;
        if (!y_ready &&
            (rc = P.run(q, WT_PROF_PERMUTE, 0, blk * (16 + cs) + n * 8.0,
                        [&] { return wt_to_elim<double, R>(r, s.iperm, s.rscale, num.y, n, nrhs, q); })))
            return rc;
        if ((rc = wt_solver_solve_block<R>(s, num, nrhs, nullptr))) return rc;
        // x = d on the first step (x starts at zero), x += d after
        const int acc = it > 0;
        if ((rc = P.run(q, WT_PROF_PERMUTE, 0, blk * ((acc ? 32 : 16) + cs) + n * 8.0,
                        [&] { return wt_from_elim<R, double>(num.y, s.iperm, x, n, nrhs, acc, q); })))
            return rc;
        if (fused) {
            if ((rc = P.run(q, WT_PROF_RESIDUAL, 8.0 * s.nnz * nrhs,
                            s.nnz * 20.0 + (n + 1) * 4.0 + n * 16.0 + blk * (32.0 + cs), [&] {
                                return wt_residual_elim<double, R>(s.rowptr, s.colidx, s.vals, x, b, s.perm, s.rscale,
                                                                   num.y, s.partial, s.norms, n, nrhs, q);
                            })))
                return rc;
            y_ready = true;
        } else if ((rc = P.run(q, WT_PROF_RESIDUAL, 8.0 * s.nnz * nrhs, s.nnz * 20.0 + (n + 1) * 4.0 + blk * 48.0, [&] {
                        return wt_residual<double>(s.rowptr, s.colidx, s.vals, x, b, s.rb, s.partial, s.norms, n,
                                                   nrhs, q);
                    }))) {
            return rc;
        }
        WT_HIP(hipMemcpyAsync(h.data(), s.norms, 2 * nrhs * sizeof(double), hipMemcpyDeviceToHost, q));
        WT_HIP(hipStreamSynchronize(q));
        P.flush();
        ++it;
        rel = 0.0;
        double absmax = 0.0;
        for (int c = 0; c < nrhs; ++c) {
            double rr = std::sqrt(h[c]), bn = std::sqrt(h[nrhs + c]);
            double v = bn > 0 ? rr / bn : rr;
            if (!(v <= rel)) rel = v;  // NaN propagates
            if (rr > absmax) absmax = rr;
        }
        if (it == 1) first = rel;
        r = s.rb;
        if (st) st->abs_residual = absmax;
        if (rel <= tol || !std::isfinite(rel)) break;
        if (tol > 0 && rel >= prev) {
            stalled = true;
            break;
        }
        prev = rel;
    }
    const bool conv = tol <= 0 || rel <= tol;
    if (st) {
        st->iterations = it;
        st->nrhs = nrhs;
        st->rel_residual = rel;
        st->rel_residual_first = first;
        st->converged = conv ? 1 : 0;
        st->stalled = stalled ? 1 : 0;
    }
    if (!std::isfinite(rel)) return WT_ERR_PIVOT;
    return conv ? WT_OK : WT_ERR_NOCONV;
}

template <typename R>
int pivot_report(const wt_numeric<R>& num, wt_pivot_report* out)
{
    if (!num.have_pivstat) return WT_ERR_STATE;
    float mr, g;
    std::memcpy(&mr, &num.pivstat[2], 4);
    std::memcpy(&g, &num.pivstat[3], 4);
    out->zero_pivots = num.pivstat[0] ? 1 : 0;
    out->tiny_pivots = num.pivstat[1];
    out->min_pivot_ratio = mr;
    out->max_growth = g;
    return WT_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

int wt_solver_create(int nx, int ny, int nz, int leaf_nodes, wt_solver** out)
{
    if (!out) return WT_ERR_ARG;
    *out = nullptr;
    wt_solver* h = new (std::nothrow) wt_solver();
    if (!h) return WT_ERR_NOMEM;
    int rc = wt_solver_init(h->s, nx, ny, nz, leaf_nodes > 0 ? leaf_nodes : 64);
    if (rc) {
        wt_solver_free(h->s);
        delete h;
        return rc;
    }
    *out = h;
    return WT_OK;
}

int wt_symbolic_info(int nx, int ny, int nz, int leaf_nodes, int64_t* info, double* flops)
{
    if (!info) return WT_ERR_ARG;
    wt_symbolic S;
    int rc = wt_symbolic_build(nx, ny, nz, leaf_nodes > 0 ? leaf_nodes : 64, WT_NB, nullptr, nullptr, S);
    if (rc) return rc;
    info[0] = S.n;
    info[1] = wt_nnz(nx, ny, nz);
    info[2] = (int64_t)S.fronts.size();
    info[3] = (int64_t)S.levels.size();
    info[4] = S.max_p;
    info[5] = S.max_u;
    info[6] = S.factor_len;
    info[7] = S.cb_len[0] + S.cb_len[1];
    if (flops) *flops = S.flops;
    return WT_OK;
}

int64_t wt_estimate_bytes(int nx, int ny, int nz, int leaf_nodes, int factor_prec, int nrhs)
{
    wt_symbolic S;
    if (wt_symbolic_build(nx, ny, nz, leaf_nodes > 0 ? leaf_nodes : 64, WT_NB, nullptr, nullptr, S) != WT_OK)
        return -1;
    const int64_t cs = factor_prec == WT_PREC_F32 ? 8 : 16, nnz = wt_nnz(nx, ny, nz);
    const int64_t rb = nrhs < 64 ? nrhs : 64;
    int64_t b = (S.n + 1) * 4 + nnz * (4 + 8 + 16) + S.n * 16;          // pattern, map, values, perm, scale
    b += (int64_t)S.upd_idx.size() * 8 + (int64_t)S.relmap.size() * 4;
    b += (S.n / 3) * 40;                                                // model
    b += (S.factor_len + S.cb_len[0] + S.cb_len[1] + S.tile_len) * cs + S.piv_len * 4;
    b += (S.n + S.w_len[0] + S.w_len[1]) * rb * cs;                    // solve blocks
    b += 3 * S.n * rb * 16;                                             // refinement blocks
    return b;
}

int wt_solver_destroy(wt_solver* h)
{
    if (!h) return WT_OK;
    wt_solver_free(h->s);
    delete h;
    return WT_OK;
}

int wt_solver_info(const wt_solver* h, int64_t* info, double* flops)
{
    if (!h || !info) return WT_ERR_ARG;
    const wt_symbolic& S = h->s.S;
    info[0] = S.n;
    info[1] = h->s.nnz;
    info[2] = (int64_t)S.fronts.size();
    info[3] = (int64_t)S.levels.size();
    info[4] = S.max_p;
    info[5] = S.max_u;
    info[6] = S.factor_len;
    info[7] = S.cb_len[0] + S.cb_len[1];
    if (flops) *flops = S.flops;
    return WT_OK;
}

int64_t wt_solver_bytes(const wt_solver* h, int factor_prec, int nrhs)
{
    if (!h) return -1;
    const wt_symbolic& S = h->s.S;
    const int64_t cs = factor_prec == WT_PREC_F32 ? 8 : 16;
    int64_t b = 0;
    b += (S.n + 1) * 4 + h->s.nnz * (4 + 8 + 16);                   // pattern, scatter map, fp64 values
    b += S.n * 8 + (int64_t)S.upd_idx.size() * 8 + (int64_t)S.relmap.size() * 4;
    b += (int64_t)S.fronts.size() * (int64_t)sizeof(wt_dfront) + 3 * S.n / 3 * 40;  // model
    b += (S.factor_len + S.cb_len[0] + S.cb_len[1] + S.tile_len) * cs + S.piv_len * 4;
    b += (S.n + S.w_len[0] + S.w_len[1]) * nrhs * cs;                 // solve blocks
    b += 3 * S.n * (int64_t)(nrhs < 64 ? nrhs : 64) * 16;             // refinement blocks
    return b;
}

int wt_solver_set_model(wt_solver* h, const void* lam, const void* mu, const double* rho, int on_device)
{
    if (!h || !lam || !mu || !rho) return WT_ERR_ARG;
    wt_solver_impl& s = h->s;
    const int64_t nodes = s.n / 3;
    int rc;
    if (!s.lam && ((rc = dalloc(&s.lam, nodes)) || (rc = dalloc(&s.mu, nodes)) || (rc = dalloc(&s.rho, nodes))))
        return rc;
    hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    WT_HIP(hipMemcpyAsync(s.lam, lam, nodes * 16, k, s.stream));
    WT_HIP(hipMemcpyAsync(s.mu, mu, nodes * 16, k, s.stream));
    WT_HIP(hipMemcpyAsync(s.rho, rho, nodes * 8, k, s.stream));
    WT_HIP(hipStreamSynchronize(s.stream));
    return WT_OK;
}

namespace {
// 10 epsilon of the working precision, the reference's threshold on omega
// (dispersion.f:43, EPSILON(1.0)*10 in its REAL)
constexpr double kOmegaEps = 10.0 * 2.220446049250313e-16;

int dispersion_dev(int64_t nodes, const double* vp, const double* vs, const double* qpi, const double* qsi,
                   const double* rho, double freqbase, double omega_re, cpx<double>* lam, cpx<double>* mu,
                   hipStream_t st)
{
    const double pi = 3.141592653589793;
    const double omegabase = freqbase * 2.0 * pi;
    const int att = (omega_re > kOmegaEps && omegabase > 0.0) ? 1 : 0;
    const double fac1 = att ? std::log(omega_re / omegabase) : 0.0;
    return wt_launch_dispersion(nodes, vp, vs, qpi, qsi, rho, fac1, att, lam, mu, st);
}
}  // namespace

int wt_dispersion(int64_t nodes, const double* vp, const double* vs, const double* qpi, const double* qsi,
                  const double* rho, double freqbase, double omega_re, void* lam, void* mu, int on_device,
                  void* stream)
{
    if (nodes < 1 || !vp || !vs || !qpi || !qsi || !rho || !lam || !mu) return WT_ERR_ARG;
    const hipStream_t st = (hipStream_t)stream;
    wt_rt_stream_guard g(stream);
    if (on_device) {
        int rc = dispersion_dev(nodes, vp, vs, qpi, qsi, rho, freqbase, omega_re, (cpx<double>*)lam, (cpx<double>*)mu,
                                st);
        if (rc) return rc;
        WT_HIP(hipStreamSynchronize(st));
        return WT_OK;
    }
    double* d = nullptr;
    if (rt_malloc_async((void**)&d, (size_t)nodes * (5 * 8 + 32), st) != hipSuccess) return WT_ERR_NOMEM;
    const double* src[5] = {vp, vs, qpi, qsi, rho};
    int rc = WT_OK;
    for (int k = 0; k < 5 && !rc; ++k)
        if (hipMemcpyAsync(d + k * nodes, src[k], nodes * 8, hipMemcpyHostToDevice, st) != hipSuccess) rc = WT_ERR_HIP;
    cpx<double>* dl = (cpx<double>*)(d + 5 * nodes);
    cpx<double>* dm = dl + nodes;
    if (!rc) rc = dispersion_dev(nodes, d, d + nodes, d + 2 * nodes, d + 3 * nodes, d + 4 * nodes, freqbase, omega_re,
                                 dl, dm, st);
    if (!rc && (hipMemcpyAsync(lam, dl, nodes * 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(mu, dm, nodes * 16, hipMemcpyDeviceToHost, st) != hipSuccess))
        rc = WT_ERR_HIP;
    (void)rt_free_async(d, st);
    if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = WT_ERR_HIP;
    return rc;
}

int wt_solver_set_model_q(wt_solver* h, const double* vp, const double* vs, const double* qpi, const double* qsi,
                          const double* rho, double freqbase, double omega_re)
{
    if (!h || !vp || !vs || !qpi || !qsi || !rho) return WT_ERR_ARG;
    wt_solver_impl& s = h->s;
    const int64_t nodes = s.n / 3;
    int rc;
    if (!s.lam && ((rc = dalloc(&s.lam, nodes)) || (rc = dalloc(&s.mu, nodes)) || (rc = dalloc(&s.rho, nodes))))
        return rc;
    double* d = nullptr;
    if (rt_malloc_async((void**)&d, (size_t)nodes * 4 * 8, s.stream) != hipSuccess) return WT_ERR_NOMEM;
    const double* src[4] = {vp, vs, qpi, qsi};
    rc = WT_OK;
    for (int k = 0; k < 4 && !rc; ++k)
        if (hipMemcpyAsync(d + k * nodes, src[k], nodes * 8, hipMemcpyHostToDevice, s.stream) != hipSuccess)
            rc = WT_ERR_HIP;
    if (!rc && hipMemcpyAsync(s.rho, rho, nodes * 8, hipMemcpyHostToDevice, s.stream) != hipSuccess) rc = WT_ERR_HIP;
    if (!rc) rc = dispersion_dev(nodes, d, d + nodes, d + 2 * nodes, d + 3 * nodes, s.rho, freqbase, omega_re, s.lam,
                                 s.mu, s.stream);
    (void)rt_free_async(d, s.stream);
    if (hipStreamSynchronize(s.stream) != hipSuccess && !rc) rc = WT_ERR_HIP;
    return rc;
}

int wt_solver_assemble(wt_solver* h, const wt_grid_params* p)
{
    if (!h || !p) return WT_ERR_ARG;
    wt_solver_impl& s = h->s;
    if (!s.lam) return WT_ERR_STATE;
    if (p->nx != s.S.nx || p->ny != s.S.ny || p->nz != s.S.nz) return WT_ERR_ARG;
    int rc;
    if (!s.vals && (rc = dalloc(&s.vals, s.nnz))) return rc;
    WT_HIP(hipMemsetAsync(s.d_err, 0, sizeof(int), s.stream));
    const double nodes = (double)(s.n / 3);
    if ((rc = s.prof.run(s.stream, WT_PROF_ASSEMBLE, 0, s.nnz * 16.0 + nodes * 40.0, [&] {
             return wt_launch_values<double>(*p, s.lam, s.mu, s.rho, s.vals, s.d_err, s.stream);
         })))
        return rc;
    int he = 0;
    WT_HIP(hipMemcpyAsync(&he, s.d_err, sizeof(int), hipMemcpyDeviceToHost, s.stream));
    WT_HIP(hipStreamSynchronize(s.stream));
    s.prof.flush();
    if (he) return WT_ERR_ARG;
    s.have_vals = true;
    s.last_p = *p;
    s.have_p = true;
    return WT_OK;
}

int wt_solver_set_values(wt_solver* h, const void* vals, int prec, int on_device)
{
    if (!h || !vals) return WT_ERR_ARG;
    wt_solver_impl& s = h->s;
    s.have_p = false;
    int rc;
    if (!s.vals && (rc = dalloc(&s.vals, s.nnz))) return rc;
    hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (prec == WT_PREC_F64) {
        WT_HIP(hipMemcpyAsync(s.vals, vals, s.nnz * 16, k, s.stream));
    } else if (prec == WT_PREC_F32) {
        cpx<float>* t = nullptr;
        if ((rc = dalloc(&t, s.nnz))) return rc;
        rc = hipMemcpyAsync(t, vals, s.nnz * 8, k, s.stream) == hipSuccess ? WT_OK : WT_ERR_HIP;
        if (!rc) rc = wt_widen_c64(t, s.vals, s.nnz, s.stream);
        if (hipStreamSynchronize(s.stream) != hipSuccess && !rc) rc = WT_ERR_HIP;
        (void)rt_free(t);
        if (rc) return rc;
    } else {
        return WT_ERR_ARG;
    }
    WT_HIP(hipStreamSynchronize(s.stream));
    s.have_vals = true;
    return WT_OK;
}

int wt_solver_get_values(wt_solver* h, void* vals)
{
    if (!h || !vals || !h->s.have_vals) return WT_ERR_ARG;
    WT_HIP(rt_memcpy(vals, h->s.vals, h->s.nnz * 16, hipMemcpyDeviceToHost));
    return WT_OK;
}

int wt_solver_factor(wt_solver* h, int factor_prec)
{
    if (!h) return WT_ERR_ARG;
    if (factor_prec == WT_PREC_F32) return wt_solver_factor_run<float>(h->s, h->s.f32);
    if (factor_prec == WT_PREC_F64) return wt_solver_factor_run<double>(h->s, h->s.f64);
    return WT_ERR_ARG;
}

int wt_solver_pivot_report(const wt_solver* h, int factor_prec, wt_pivot_report* out)
{
    if (!h || !out) return WT_ERR_ARG;
    if (factor_prec == WT_PREC_F32) return pivot_report(h->s.f32, out);
    if (factor_prec == WT_PREC_F64) return pivot_report(h->s.f64, out);
    return WT_ERR_ARG;
}

int wt_solver_autotune(wt_solver* h, int factor_prec, double* speedup)
{
    if (!h) return WT_ERR_ARG;
    if (factor_prec == WT_PREC_F32) return wt_solver_autotune<float>(h->s, h->s.f32, speedup);
    if (factor_prec == WT_PREC_F64) return wt_solver_autotune<double>(h->s, h->s.f64, speedup);
    return WT_ERR_ARG;
}

int wt_solver_set_graphs(wt_solver* h, int enable)
{
    if (!h) return WT_ERR_ARG;
    h->s.use_graphs = enable != 0;
    return WT_OK;
}

int wt_solver_set_splitk(wt_solver* h, int min_depth)
{
    if (!h || min_depth < 0) return WT_ERR_ARG;
    wt_solver_impl& s = h->s;
    s.splitk_depth = min_depth;
    // the solve plans are rebuilt (same launches, so stored tile tables stay
    // valid) at the next solve
    s.f32.snrhs = 0;
    s.f64.snrhs = 0;
    s.f32.drop_graphs();
    s.f64.drop_graphs();
    return WT_OK;
}

int wt_solver_set_gemm_split(wt_solver* h, int enable)
{
    if (!h) return WT_ERR_ARG;
    wt_solver_impl& s = h->s;
    s.gemm_split = enable != 0;
    for (wt_gemm_plan<float>* pl : {&s.f32.fplan, &s.f32.splan}) {
        pl->allow_split = s.gemm_split;
        for (auto& L : pl->launches) {
            if (!s.gemm_split && L.var == WT_GEMM_SPLIT) L.var = 0;
            else if (!s.gemm_split && L.var == WT_GEMM_SPLIT64) L.var = 2;
            else if (s.gemm_split && L.var == 0) L.var = WT_GEMM_SPLIT;
        }
    }
    s.f32.drop_graphs();
    return WT_OK;
}

int wt_solver_get_tiles(wt_solver* h, int factor_prec, int which, int nrhs, int32_t* vars, int64_t* count)
{
    if (!h || !count) return WT_ERR_ARG;
    if (factor_prec == WT_PREC_F32)
        return wt_solver_plan_tiles<float>(h->s, h->s.f32, which, nrhs, vars, nullptr, count);
    if (factor_prec == WT_PREC_F64)
        return wt_solver_plan_tiles<double>(h->s, h->s.f64, which, nrhs, vars, nullptr, count);
    return WT_ERR_ARG;
}

int wt_solver_set_tiles(wt_solver* h, int factor_prec, int which, int nrhs, const int32_t* vars, int64_t count)
{
    if (!h || !vars) return WT_ERR_ARG;
    if (factor_prec == WT_PREC_F32)
        return wt_solver_plan_tiles<float>(h->s, h->s.f32, which, nrhs, nullptr, vars, &count);
    if (factor_prec == WT_PREC_F64)
        return wt_solver_plan_tiles<double>(h->s, h->s.f64, which, nrhs, nullptr, vars, &count);
    return WT_ERR_ARG;
}

int wt_solver_apply(wt_solver* h, int nrhs, const void* x, void* y, int on_device)
{
    if (!h || !x || !y || nrhs < 1) return WT_ERR_ARG;
    wt_solver_impl& s = h->s;
    if (!s.have_vals) return WT_ERR_STATE;
    const int chunk = nrhs < 64 ? nrhs : 64;
    int rc;
    if ((rc = ensure_ir(s, chunk))) return rc;
    hipMemcpyKind kin = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    hipMemcpyKind kout = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    const cpx<double>* xs = (const cpx<double>*)x;
    cpx<double>* yd = (cpx<double>*)y;
    for (int c0 = 0; c0 < nrhs; c0 += chunk) {
        const int nc = nrhs - c0 < chunk ? nrhs - c0 : chunk;
        if (nc == nrhs) WT_HIP(hipMemcpyAsync(s.xb, xs, s.n * nc * 16, kin, s.stream));
        else WT_HIP(hipMemcpy2DAsync(s.xb, nc * 16, xs + c0, (size_t)nrhs * 16, nc * 16, s.n, kin, s.stream));
        if ((rc = wt_residual<double>(s.rowptr, s.colidx, s.vals, s.xb, nullptr, s.rb, s.partial, s.norms, s.n, nc,
                                      s.stream)))
            return rc;
        if (nc == nrhs) WT_HIP(hipMemcpyAsync(yd, s.rb, s.n * nc * 16, kout, s.stream));
        else WT_HIP(hipMemcpy2DAsync(yd + c0, (size_t)nrhs * 16, s.rb, nc * 16, nc * 16, s.n, kout, s.stream));
        WT_HIP(hipStreamSynchronize(s.stream));
    }
    return WT_OK;
}

int wt_solver_planewaves(wt_solver* h, int nrhs, const double* dirs, double kmag, double dx, void* d_x)
{
    if (!h || !dirs || !d_x || nrhs < 1) return WT_ERR_ARG;
    wt_solver_impl& s = h->s;
    double* dd = nullptr;
    if (rt_malloc(&dd, 3 * nrhs * sizeof(double)) != hipSuccess) return WT_ERR_NOMEM;
    int rc = hipMemcpyAsync(dd, dirs, 3 * nrhs * sizeof(double), hipMemcpyHostToDevice, s.stream) == hipSuccess
                 ? WT_OK : WT_ERR_HIP;
    if (!rc) rc = wt_planewave((cpx<double>*)d_x, s.S.nx, s.S.ny, s.S.nz, dx, kmag, dd, nrhs, s.stream);
    if (hipStreamSynchronize(s.stream) != hipSuccess && !rc) rc = WT_ERR_HIP;
    (void)rt_free(dd);
    return rc;
}

int wt_solver_stream(const wt_solver* h, void** stream)
{
    if (!h || !stream) return WT_ERR_ARG;
    *stream = (void*)h->s.stream;
    return WT_OK;
}

int wt_solver_scattered_source(wt_solver* h, const void* lam_bg, const void* mu_bg, const double* rho_bg,
                               int on_device, int nrhs, const void* d_ubg, const double* srcfn, void* d_sfld)
{
    if (!h || !lam_bg || !mu_bg || !rho_bg || !d_ubg || !d_sfld || nrhs < 1) return WT_ERR_ARG;
    wt_solver_impl& s = h->s;
    if (!s.have_vals || !s.have_p) return WT_ERR_STATE;
    const int64_t nodes = s.n / 3;
    cpx<double>*lb = nullptr, *mb = nullptr, *ab = nullptr, *sf = nullptr;
    double* rb = nullptr;
    int rc = WT_OK;
    // scratch is stream-ordered: released after the work queued on the
    // handle's stream, without waiting for other streams
    auto cleanup = [&] {
        for (void* p : {(void*)lb, (void*)mb, (void*)rb, (void*)ab, (void*)sf})
            if (p) (void)rt_free_async(p, s.stream);
        (void)hipStreamSynchronize(s.stream);
    };
    if ((rc = salloc(&lb, nodes, s.stream)) || (rc = salloc(&mb, nodes, s.stream)) ||
        (rc = salloc(&rb, nodes, s.stream)) || (rc = salloc(&ab, s.nnz, s.stream)) ||
        (srcfn && (rc = salloc(&sf, nrhs, s.stream)))) {
        cleanup();
        return rc;
    }
    const hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (hipMemcpyAsync(lb, lam_bg, nodes * 16, k, s.stream) != hipSuccess ||
        hipMemcpyAsync(mb, mu_bg, nodes * 16, k, s.stream) != hipSuccess ||
        hipMemcpyAsync(rb, rho_bg, nodes * 8, k, s.stream) != hipSuccess ||
        (sf && hipMemcpyAsync(sf, srcfn, nrhs * 16, hipMemcpyHostToDevice, s.stream) != hipSuccess) ||
        hipMemsetAsync(s.d_err, 0, sizeof(int), s.stream) != hipSuccess) {
        cleanup();
        return WT_ERR_HIP;
    }
    // the background operator at the parameters of the last assembly
    rc = wt_launch_values<double>(s.last_p, lb, mb, rb, ab, s.d_err, s.stream);
    const cpx<double>* u = (const cpx<double>*)d_ubg;
    cpx<double>* out = (cpx<double>*)d_sfld;
    // column groups of at most 64 (one lane per column), strided in the block
    for (int c0 = 0; !rc && c0 < nrhs; c0 += 64) {
        const int nc = nrhs - c0 < 64 ? nrhs - c0 : 64;
        rc = wt_launch_scatsrc(s.rowptr, s.colidx, s.vals, ab, u + c0, sf ? sf + c0 : nullptr, out + c0, s.n, nc,
                               s.stream, nrhs, nrhs);
    }
    int he = 0;
    if (!rc && (hipMemcpyAsync(&he, s.d_err, sizeof(int), hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
                hipStreamSynchronize(s.stream) != hipSuccess))
        rc = WT_ERR_HIP;
    if (rc == WT_OK && he) rc = WT_ERR_ARG;
    if (rc != WT_OK) (void)hipStreamSynchronize(s.stream);
    cleanup();
    return rc;
}

int wt_dev_copy(void* d_dst, const void* d_src, size_t bytes)
{
    if (!d_dst || !d_src || bytes % 16) return WT_ERR_ARG;
    wt_rt_shared g;  // null-stream launch
    int rc = wt_copy<double>((const cpx<double>*)d_src, (cpx<double>*)d_dst, (int64_t)(bytes / 16), nullptr);
    if (rc) return rc;
    WT_HIP(rt_device_sync());
    return WT_OK;
}

int wt_solver_profile(wt_solver* h, int enable)
{
    if (!h) return WT_ERR_ARG;
    WT_HIP(hipStreamSynchronize(h->s.stream));
    h->s.prof.flush();
    h->s.prof.reset();
    h->s.prof.on = enable != 0;
    return WT_OK;
}

int wt_solver_profile_read(wt_solver* h, int cls, int64_t* launches, double* ms, double* flops, double* bytes)
{
    if (!h || cls < 0 || cls >= WT_PROF_NCLASS) return WT_ERR_ARG;
    const wt_prof& P = h->s.prof;
    if (launches) *launches = P.launches[cls];
    if (ms) *ms = P.ms[cls];
    if (flops) *flops = P.flops[cls];
    if (bytes) *bytes = P.bytes[cls];
    return WT_OK;
}

int wt_solver_release(wt_solver* h, int factor_prec)
{
    if (!h) return WT_ERR_ARG;
    if (factor_prec == WT_PREC_F32) h->s.f32.release();
    else if (factor_prec == WT_PREC_F64) h->s.f64.release();
    else return WT_ERR_ARG;
    return WT_OK;
}

int wt_solver_solve(wt_solver* h, int factor_prec, int nrhs, const void* b, void* x, int on_device, double tol,
                    int max_iter, wt_solve_stats* st)
{
    if (!h || !b || !x || nrhs < 1 || max_iter < 1) return WT_ERR_ARG;
    wt_solver_impl& s = h->s;
    if (!s.have_vals) return WT_ERR_STATE;
    bool f32 = factor_prec == WT_PREC_F32;
    if (!f32 && factor_prec != WT_PREC_F64) return WT_ERR_ARG;
    if ((f32 && !s.f32.ready) || (!f32 && !s.f64.ready)) return WT_ERR_STATE;
    const int chunk = nrhs < 64 ? nrhs : 64;
    int rc;
    if ((rc = ensure_ir(s, chunk))) return rc;
    wt_solve_stats agg{};
    agg.rel_residual = 0;
    agg.rel_residual_first = 0;
    agg.converged = 1;
    int result = WT_OK;
    auto t0 = std::chrono::steady_clock::now();
    for (int c0 = 0; c0 < nrhs; c0 += chunk) {
        const int nc = nrhs - c0 < chunk ? nrhs - c0 : chunk;
        // gather this chunk's columns (row-major n x nrhs) into a dense n x nc block
        const cpx<double>* src = (const cpx<double>*)b;
        cpx<double>* dstx = (cpx<double>*)x;
        const cpx<double>* bblk;
        cpx<double>* xblk;
        if (on_device && nc == nrhs) {
            bblk = src;
            xblk = dstx;
        } else {
            hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
            if (nc == nrhs) WT_HIP(hipMemcpyAsync(s.bb, src, s.n * nc * 16, k, s.stream));
            else WT_HIP(hipMemcpy2DAsync(s.bb, nc * 16, src + c0, (size_t)nrhs * 16, nc * 16, s.n, k, s.stream));
            bblk = s.bb;
            xblk = s.xb;
        }
        wt_solve_stats cs{};
        rc = f32 ? refine<float>(s, s.f32, nc, bblk, xblk, tol, max_iter, &cs)
                 : refine<double>(s, s.f64, nc, bblk, xblk, tol, max_iter, &cs);
        if (rc == WT_ERR_NOCONV) result = rc;  // keep going: x and stats are still delivered
        else if (rc) return rc;
        if (!cs.converged) agg.converged = 0;
        if (cs.stalled) agg.stalled = 1;
        if (xblk == s.xb) {
            hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
            if (nc == nrhs) WT_HIP(hipMemcpyAsync(dstx, s.xb, s.n * nc * 16, k, s.stream));
            else WT_HIP(hipMemcpy2DAsync(dstx + c0, (size_t)nrhs * 16, s.xb, nc * 16, nc * 16, s.n, k, s.stream));
        }
        WT_HIP(hipStreamSynchronize(s.stream));
        agg.iterations = cs.iterations > agg.iterations ? cs.iterations : agg.iterations;
        agg.rel_residual = cs.rel_residual > agg.rel_residual ? cs.rel_residual : agg.rel_residual;
        agg.rel_residual_first =
            cs.rel_residual_first > agg.rel_residual_first ? cs.rel_residual_first : agg.rel_residual_first;
        agg.abs_residual = cs.abs_residual > agg.abs_residual ? cs.abs_residual : agg.abs_residual;
    }
    agg.nrhs = nrhs;
    agg.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (st) *st = agg;
    return result;
}

// ---------------------------------------------------------------------------
// Drop-in for strumpack_serial_finter (strumpack_finter.c:9-120): one static
// solver; job 1 = analyse the CSR pattern (1-based, left untouched), job 2 =
// numeric factorization of the complex64 values, job 3 = solve one
// complex64 right-hand side with refinement to the reference's tolerances
// (rctol 1e-4, actol 1e-10, strumpack_finter.c:39-40), job -1 = release.
// The grid is recovered from the pattern; any other pattern is rejected.

static wt_solver* g_legacy = nullptr;
static bool g_factored = false;

static int grid_from_pattern(int n, int nnz, const int* irptr, const int* jcptr, int* nx, int* ny, int* nz)
{
    if (n < 24 || n % 3 || irptr[0] != 1 || irptr[1] - irptr[0] != 24) return WT_ERR_PATTERN;
    const int z = (jcptr[6] - 1) / 3, xz = (jcptr[12] - 1) / 3;
    if (z < 2 || xz % z || xz / z < 2 || (n / 3) % xz) return WT_ERR_PATTERN;
    *nz = z;
    *nx = xz / z;
    *ny = (n / 3) / xz;
    if (*ny < 2 || wt_nnz(*nx, *ny, *nz) != nnz) return WT_ERR_PATTERN;
    std::vector<int32_t> r(n + 1), c(nnz);
    if (wt_graph(*nx, *ny, *nz, 1, r.data(), c.data()) != WT_OK) return WT_ERR_HIP;
    if (std::memcmp(r.data(), irptr, (size_t)(n + 1) * 4) || std::memcmp(c.data(), jcptr, (size_t)nnz * 4))
        return WT_ERR_PATTERN;
    return WT_OK;
}

void strumpack_serial_finter(int* job, int* nrows, int* nnz, int* irptr, int* jcptr, void* a, void* rhs, void* sol,
                             int* ierr)
{
    const char* fcnm = "strumpack_serial_finter";
    *ierr = 0;
    if (*job == 1) {
        if (g_legacy) {
            printf("%s: Error already initialized!\n", fcnm);
            *ierr = 1;
            return;
        }
        int nx, ny, nz;
        if (grid_from_pattern(*nrows, *nnz, irptr, jcptr, &nx, &ny, &nz) != WT_OK) {
            printf("%s: Error reordering matrix!\n", fcnm);
            *ierr = 1;
            return;
        }
        if (wt_solver_create(nx, ny, nz, 64, &g_legacy) != WT_OK) {
            printf("%s: Strumpack error!\n", fcnm);
            *ierr = 1;
            return;
        }
        g_factored = false;
    } else if (*job == 2) {
        g_factored = false;
        if (!g_legacy) {
            printf("%s: Error not initialized!\n", fcnm);
            *ierr = 1;
            return;
        }
        if (wt_solver_set_values(g_legacy, a, WT_PREC_F32, 0) != WT_OK ||
            wt_solver_factor(g_legacy, WT_PREC_F32) != WT_OK) {
            printf("%s: Error factoring matrix!\n", fcnm);
            *ierr = 1;
            return;
        }
        g_factored = true;
    } else if (*job == 3) {
        if (!g_legacy || !g_factored) {
            printf("%s: Error matrix not factorized!\n", fcnm);
            *ierr = 1;
            return;
        }
        const int64_t n = *nrows;
        std::vector<cpx<double>> b(n), x(n);
        const cpx<float>* r32 = (const cpx<float>*)rhs;
        for (int64_t i = 0; i < n; ++i) b[i] = cpx<double>(r32[i].re, r32[i].im);
        wt_solve_stats st{};
        int rc = wt_solver_solve(g_legacy, WT_PREC_F32, 1, b.data(), x.data(), 0, 1e-4, 10, &st);
        // converged when either the relative or the absolute criterion holds
        // (rctol, actol); a block above both is a failed solve
        if ((rc != WT_OK && rc != WT_ERR_NOCONV) || (st.rel_residual > 1e-4 && st.abs_residual > 1e-10)) {
            printf("%s: Error in solution phase1\n", fcnm);
            *ierr = 1;
            if (rc != WT_OK && rc != WT_ERR_NOCONV) return;
        }
        cpx<float>* s32 = (cpx<float>*)sol;
        for (int64_t i = 0; i < n; ++i) s32[i] = cpx<float>((float)x[i].re, (float)x[i].im);
    } else if (*job == -1) {
        // the reference driver passes already-deallocated arrays here
        // (elwave.f90:184-190): they are not touched
        wt_solver_destroy(g_legacy);
        g_legacy = nullptr;
        g_factored = false;
    }
}

int strumpack_parallel_finter(void) { return 0; }

}  // extern "C"
