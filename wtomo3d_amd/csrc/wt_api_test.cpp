// C ABI: single-problem entry to the batched complex GEMM (for unit tests
// and microbenchmarks of the front kernels' workhorse).
#include <vector>

#include "wt_gemm.h"
#include "wt_rt.h"

namespace {
template <typename R>
int run_gemm(int var, int M, int N, int K, const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
             int64_t ldc, int mode, const int32_t* bperm, int reps, double* ms)
{
    if (M < 1 || N < 1 || K < 1 || lda < M || ldc < M || reps < 1 || var < -1 || var >= WT_GEMM_NVAR)
        return WT_ERR_ARG;
    const int64_t cs = sizeof(cpx<R>);
    const int64_t na = lda * K, nbv = ldb * N, nc = ldc * N;
    cpx<R>*dA = nullptr, *dB = nullptr, *dC = nullptr;
    int32_t* dp = nullptr;
    wt_gemm_prob<R>* dprob = nullptr;
    int64_t* dts = nullptr;
    int rc = WT_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipMalloc(&dA, na * cs) || hipMalloc(&dB, nbv * cs) || hipMalloc(&dC, nc * cs) ||
        hipMalloc(&dprob, sizeof(wt_gemm_prob<R>)) || hipMalloc(&dts, 2 * sizeof(int64_t)) ||
        (bperm && hipMalloc(&dp, K * 4))) {
        rc = WT_ERR_NOMEM;
    } else {
        wt_gemm_prob<R> p;
        p.A = dA; p.B = dB; p.C = dC; p.lda = lda; p.ldb = ldb; p.ldc = ldc;
        p.M = M; p.N = N; p.K = K; p.mode = mode; p.bperm = dp; p.pad = 0;
        if (var < 0) var = (sizeof(R) == 4 && N > 64 && wt_gemm_tiles<R>(M, N, 0) >= 1024) ? 0 : 2;
        int64_t ts[2] = {0, wt_gemm_tiles<R>(M, N, var)};
        if (hipMemcpy(dA, A, na * cs, hipMemcpyHostToDevice) || hipMemcpy(dB, B, nbv * cs, hipMemcpyHostToDevice) ||
            (bperm && hipMemcpy(dp, bperm, K * 4, hipMemcpyHostToDevice)) ||
            hipMemcpy(dprob, &p, sizeof(p), hipMemcpyHostToDevice) ||
            hipMemcpy(dts, ts, sizeof(ts), hipMemcpyHostToDevice) || hipEventCreate(&e0) || hipEventCreate(&e1)) {
            rc = WT_ERR_HIP;
        } else {
            for (int r = 0; r < reps && rc == WT_OK; ++r) {
                if (hipMemcpy(dC, C, nc * cs, hipMemcpyHostToDevice)) { rc = WT_ERR_HIP; break; }
                (void)hipEventRecord(e0, 0);
                rc = wt_gemm_launch<R>(dprob, dts, 1, ts[1], var, 0, mode == WT_GEMM_SUB || mode == WT_GEMM_ADD);
                (void)hipEventRecord(e1, 0);
            }
            float t = 0;
            if (rc == WT_OK && (hipEventSynchronize(e1) || hipEventElapsedTime(&t, e0, e1))) rc = WT_ERR_HIP;
            if (ms) *ms = t;
            if (rc == WT_OK && hipMemcpy(C, dC, nc * cs, hipMemcpyDeviceToHost)) rc = WT_ERR_HIP;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC); (void)hipFree(dp); (void)hipFree(dprob); (void)hipFree(dts);
    return rc;
}
}  // namespace

// var: tile variant (0..WT_GEMM_NVAR-1), or -1 for the default rule
extern "C" int wt_gemm_host_var(int prec, int var, int M, int N, int K, const void* A, int64_t lda, const void* B,
                                int64_t ldb, void* C, int64_t ldc, int mode, const int32_t* bperm, int reps,
                                double* ms)
{
    if (mode < WT_GEMM_SUB || mode > WT_GEMM_NSET) return WT_ERR_ARG;
    wt_rt_shared g;  // synchronous copies and null-stream launches
    if (prec == WT_PREC_F32) return run_gemm<float>(var, M, N, K, A, lda, B, ldb, C, ldc, mode, bperm, reps, ms);
    if (prec == WT_PREC_F64) return run_gemm<double>(var, M, N, K, A, lda, B, ldb, C, ldc, mode, bperm, reps, ms);
    return WT_ERR_ARG;
}

extern "C" int wt_gemm_host(int prec, int M, int N, int K, const void* A, int64_t lda, const void* B, int64_t ldb,
                            void* C, int64_t ldc, int mode, const int32_t* bperm, int reps, double* ms)
{
    return wt_gemm_host_var(prec, -1, M, N, K, A, lda, B, ldb, C, ldc, mode, bperm, reps, ms);
}
