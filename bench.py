"""Headline benchmark: RHS solves/s of the elastic forward solve on one MI355X.

Workload (the heterogeneous 96x96x64 configuration): dx = 200 m, a seeded
6-layer background (Vp 4-8 km/s increasing with depth, Vs = Vp/1.73,
Gardner density) plus a Gaussian random field (5 % std, 2 km correlation)
on Vp and Vs with the density perturbed by 0.3x; 10-node PML (pmld = 10 dx,
ipml = 11) with a free surface; 0.3 Hz; 64 plane-wave right-hand sides as
one block.  Before the warm-up the factorization's GEMM tile shapes are
autotuned: the package's tuning table for the workload is applied when it
exists (``wtomo3d_amd/tuning/``), otherwise one factorization per tile
variant is measured (``--retune`` forces that, ``--no-autotune`` keeps the
defaults).  One step = FP64 assembly of the operator at the frequency +
complex64 multifrontal factorization + block solve with FP64 iterative
refinement to max_j ||b_j - A x_j|| / ||b_j|| <= 1e-8.  The model, the CSR
pattern, the symbolic analysis and the right-hand sides are resident in HBM
before the timed region.  The plain complex128 path (factor + one solve) is
timed after it on the same inputs.

Right-hand sides: b_j = A u_j with u_j a homogeneous P plane wave (incidence
10-35 deg, azimuth 0-360 deg, seeded), a synthetic source with the
operator's scaling.

After the headline: the layered 64x64x48 sweep over 8 frequencies with 32
Haskell plane waves each (``sweep``, ``--no-sweep`` skips it), and the
batched small-model workload (256 models of uneven size, each its own
frequency and 1-16 right-hand sides; ``batched``, ``--no-batch``).

Prints one JSON line.  ``--gpus`` above 1 is refused: the library runs on
one GPU (frequencies, models and right-hand-side blocks are independent and
would shard across GPUs without communication).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 matrix
PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0  # dense bf16 matrix
PEAK_HBM_GBS = 8000.0
METRIC = "RHS solves/s at 96x96x64, FP32-factor+IR vs FP64 @1e-8 resid; p99 latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=96)
    ap.add_argument("--ny", type=int, default=96)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain complex128 path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-latency sample")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched small-model sample")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default GEMM tile shapes (no tuning table, no tuning factorizations)")
    ap.add_argument("--retune", action="store_true",
                    help="measure the tile shapes even when the package holds a tuning table for this workload")
    ap.add_argument("--save-tuning", metavar="PATH", default=None,
                    help="write the tuned tile shapes (factorization and block solve) to PATH as JSON")
    ap.add_argument("--no-split", action="store_true",
                    help="keep every complex64 product on the fp32 matrix cores (no bf16-split GEMMs)")
    ap.add_argument("--no-sweep", action="store_true", help="skip the multi-frequency layered sample")
    ap.add_argument("--serve-rate", type=float, default=0.7,
                    help="requests/s of the serving sample (default: the highest rate of the mix one GPU "
                         "was measured to sustain, so p50/p99 are steady-state latencies)")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="arrival window of the serving sample")
    ap.add_argument("--burst-seconds", type=float, default=3.0,
                    help="arrival window of the 20 requests/s serving burst (0: none)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel 5 times on BYTES (HBM counter calibration)")
    return ap.parse_args()


def cpu_baseline(nx, ny, nz, nrhs, budget, solves=4):
    """The oracle's dense front factorization (LAPACK via numpy/scipy) timed
    on the host on the largest fronts of the same elimination tree, within
    `budget` seconds; the tree's total work at that rate -- the
    factorization plus `solves` block solves of `nrhs` columns (8 flops per
    factor entry and column each, the same count as the GPU step's) --
    gives RHS/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mf_oracle as MF
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = min(cores, 16)
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(cores)
    except Exception:
        limiter = None
    fronts = MF.nd_fronts(nx, ny, nz)
    fl, t, sample = MF.cpu_rate_sample(fronts, np.complex64, budget)
    if limiter is not None:
        limiter.unregister() if hasattr(limiter, "unregister") else None
    rate = fl / t
    total = MF.tree_flops(fronts)
    entries = float(sum(p * p + 2 * p * u for p, u, _ in fronts))  # [L11\U11; L21] and U12
    solve_fl = solves * 8.0 * nrhs * entries
    t_tree = (total + solve_fl) / rate
    return {"value": nrhs / t_tree, "unit": "RHS/s", "cores": cores, "kind": "port",
            "sample": "complex64 dense partial LU (LAPACK getrf/trsm + gemm) of the %d largest fronts "
                      "(p,u)=%s of the %dx%dx%d elimination tree, %.1f s at %.3g GFLOP/s; the whole tree's "
                      "%.3g flop plus %d block solves of %d columns (%.3g flop) extrapolated at that rate to "
                      "%.1f s per step"
                      % (len(sample), sample[:3], nx, ny, nz, t, rate / 1e9, total, solves, nrhs, solve_fl,
                         t_tree)}


def reference_assembly(nx, ny, nz, dx, freq, vp, vs, rho, gpu_ms):
    """The reference's own assembly (SREL3D + SPARSE_ADJ2CRS, then ASMBLE with
    its SETCOEFS3D / FDFD3D / PML3D physics), compiled from its Fortran
    sources into oracle/_ref/libwtref32.so as shipped (single precision,
    one thread), timed on the host on the same model; None when that build
    is absent."""
    import ctypes
    so = os.path.join(ROOT, "oracle", "_ref", "libwtref32.so")
    if not os.path.exists(so):
        return None
    L = ctypes.CDLL(so)
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    F = lambda v: ctypes.byref(ctypes.c_float(v))
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    nv, nzr, ierr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(-1)
    t0 = time.perf_counter()
    L.wtref_graph_size(I(nx), I(ny), I(nz), ctypes.byref(nv), ctypes.byref(nzr))
    ir = np.zeros(nv.value + 1, np.int32)
    jc = np.zeros(nzr.value, np.int32)
    L.wtref_graph(I(nx), I(ny), I(nz), P(ir), P(jc))
    t_graph = time.perf_counter() - t0
    vals = np.zeros(jc.size, np.complex64)
    v32 = [np.ascontiguousarray(x, np.float32) for x in (vp, vs, rho)]
    t0 = time.perf_counter()
    L.wtref_assemble(I(nx), I(ny), I(nz), F(dx), F(freq), F(999.999), I(1), I(1), F(10 * dx), F(1e-3),
                     P(v32[0]), P(v32[1]), P(v32[2]), I(jc.size), P(ir), P(jc), P(vals), ctypes.byref(ierr))
    t_asm = time.perf_counter() - t0
    if ierr.value != 0:
        return None
    return {"s": t_asm, "graph_s": t_graph, "unit": "s per assembly", "cores": 1, "kind": "reference",
            "gpu_ms": gpu_ms, "gpu_speedup": t_asm / (gpu_ms * 1e-3) if gpu_ms else None,
            "sample": "ASMBLE of the %dx%dx%d model in complex64 (the reference as shipped, %d values), one "
                      "thread; the GPU figure is the FP64 assembly" % (nx, ny, nz, jc.size)}


GEMM_KERNEL = "gemm_batched_kernel<float, 0>"
SPLIT_KERNEL = "gemm_split_kernel<128>"


def gemm_traffic(kernel):
    """HBM bytes per launch of the roofline kernel from the committed counter
    summary (profiles/gemm_traffic.json, two rocprofv3 --pmc passes of this
    benchmark scaled by a copy-kernel calibration), when it covers it."""
    path = os.path.join(ROOT, "profiles", "gemm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    t = t.get(kernel) if "kernel" not in t else (t if t["kernel"] == kernel else None)
    if not t:
        return None
    return {"bytes_per_launch": t["bytes_per_dispatch"], "source": t.get("source", path)}


def gemm_roofline(split, kernel, achieved, traffic, g, all_gemm, steps):
    """Roofline of the factorization's dominant GEMM kernel.  On the fp32
    matrix cores: 8 M N K per complex product against the FP32 peak (3M
    executes 3/4 of it).  As bf16 splits: the executed bf16 work, 3/4 (3M)
    x 6 split pairs of the same count, against the bf16 peak, with the
    fp32-equivalent rate beside it."""
    r = {"bound": "mfma", "unit": "TFLOP/s",
         "traffic": traffic["bytes_per_launch"] if traffic else None,
         "traffic_source": traffic["source"] if traffic else None,
         "flops_per_launch": g["flops"] / max(g["launches"], 1),
         "avg_launch_ms": g["ms"] / max(g["launches"], 1),
         "launches_per_step": g["launches"] / steps, "ms_per_step": g["ms"] / steps,
         "all_factor_gemms": all_gemm}
    if split:
        executed = achieved * 0.75 * 6
        r.update(kernel=kernel + ": the factorization's large complex64 GEMMs, each fp32 product as six bf16 "
                 "split pairs (3M, fp32 accumulate); achieved counts the executed bf16 flops",
                 achieved=executed, peak=PEAK_BF16_TFLOPS, frac=executed / PEAK_BF16_TFLOPS,
                 fp32_equivalent_tflops=achieved, fp32_equivalent_frac=achieved / PEAK_FP32_TFLOPS)
    else:
        r.update(kernel=kernel + ": the factorization's complex GEMMs on the 128x128 tile (3M, fp32 MFMA); "
                 "all_factor_gemms counts every tile variant", achieved=achieved, peak=PEAK_FP32_TFLOPS,
                 frac=achieved / PEAK_FP32_TFLOPS, executed_frac_3m=0.75 * achieved / PEAK_FP32_TFLOPS)
    return r


def progress(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def serving_sample(rate, seconds, burst_seconds):
    """Latency of the serving loop on a bounded sample of the Poisson
    request stream: heterogeneous models with sides uniform in 32-80 nodes,
    0.05-0.5 Hz, 1-64 plane-wave sources, 30 % of requests on a (model,
    frequency) whose factor is resident; realtime arrivals at `rate` for
    `seconds`, then a bounded window of arrivals at 20 requests/s (the
    configured stream rate) for `burst_seconds`, showing admission under a
    rate beyond one GPU."""
    from wtomo3d_amd import serving
    rep = serving.run_stream(rate, seconds, seed=1005)
    if burst_seconds > 0:
        rep["burst_20_per_s"] = serving.run_stream(20.0, burst_seconds, seed=1006)
    return rep


def sweep_sample():
    """The layered multi-frequency workload (wtomo3d_amd.sweep.benchmark)."""
    from wtomo3d_amd import sweep
    return sweep.benchmark()


def batch_sample():
    """The batched small-model workload (wtomo3d_amd.batch.benchmark)."""
    from wtomo3d_amd import batch
    return batch.benchmark(progress=progress)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: the library runs on one GPU only (--gpus 1); multi-GPU sharding "
              "of frequencies / models / right-hand-side blocks is not built", file=sys.stderr)
        sys.exit(2)
    import wtomo3d_amd as W
    if a.copy_probe:
        L = W.lib()
        src, dst = ctypes.c_void_p(), ctypes.c_void_p()
        W._lib.check(L.wt_dev_alloc(a.copy_probe, ctypes.byref(src)), "alloc")
        W._lib.check(L.wt_dev_alloc(a.copy_probe, ctypes.byref(dst)), "alloc")
        for _ in range(5):
            W._lib.check(L.wt_dev_copy(dst, src, a.copy_probe), "copy")
        print(json.dumps({"copy_probe_bytes": a.copy_probe, "launches": 5}))
        return
    from wtomo3d_amd import media
    from wtomo3d_amd.solver import Solver

    nx, ny, nz, nrhs, dx, freq = a.nx, a.ny, a.nz, a.nrhs, 200.0, 0.3
    omega = 2 * np.pi * freq
    vp, vs, rho = media.heterogeneous(nx, ny, nz, dx, seed=1003)
    lam, mu, rho1 = media.lame(vp, vs, rho)
    t0 = time.perf_counter()
    s = Solver(nx, ny, nz)
    t_symbolic = time.perf_counter() - t0
    info = s.info()
    s.set_model(lam, mu, rho1)
    g = W.grid_params(nx, ny, nz, dx, omega, pmld=10 * dx)
    s.assemble(g)
    U = media.plane_wave_fields(nx, ny, nz, dx, omega, float(np.median(vp)), nrhs, seed=2003)
    B = s.apply(U)
    L = W.lib()
    n = s.n
    dB, dX = ctypes.c_void_p(), ctypes.c_void_p()
    W._lib.check(L.wt_dev_alloc(n * nrhs * 16, ctypes.byref(dB)), "alloc")
    W._lib.check(L.wt_dev_alloc(n * nrhs * 16, ctypes.byref(dX)), "alloc")
    W._lib.check(L.wt_h2d(dB, B.ctypes.data_as(ctypes.c_void_p), B.nbytes), "h2d")
    del U
    st = W._lib.WtSolveStats()

    def step(prec, tol, maxit):
        W._lib.check(L.wt_solver_assemble(s._h, ctypes.byref(g)), "assemble")
        W._lib.check(L.wt_solver_factor(s._h, prec), "factor")
        W._lib.check(L.wt_solver_solve(s._h, prec, nrhs, dB, dX, 1, tol, maxit, ctypes.byref(st)), "solve")
        return st.rel_residual, st.iterations

    if a.no_split:
        s.set_gemm_split(False)
    tune = None
    table = os.path.join(ROOT, "wtomo3d_amd", "tuning", "fp32_%dx%dx%d_nrhs%d%s.json"
                         % (nx, ny, nz, nrhs, "_nosplit" if a.no_split else ""))
    if not a.no_autotune and not a.retune and os.path.exists(table):
        with open(table) as f:
            tt = json.load(f)
        try:
            s.set_tiles(tt["factor"], 32, 0)
            s.set_tiles(tt["solve"], 32, 1, nrhs)
            tune = {"source": os.path.relpath(table, ROOT), "gemm_speedup_when_measured": tt.get("gemm_speedup"),
                    "note": "tile shapes measured by the autotuner on MI355X and stored in the package"}
        except RuntimeError:
            tune = None  # the plans changed since the table was written: measure again
    if not a.no_autotune and tune is None:
        progress("autotuning the factorization's GEMM tile shapes")
        t0 = time.perf_counter()
        sp = s.autotune(32)
        tune = {"gemm_speedup": sp, "s": time.perf_counter() - t0,
                "note": "one factorization per tile variant before the warm-up (the block solve's tile shapes "
                        "when its plan is built); every GEMM launch keeps its fastest variant"}
    progress("symbolic %.2f s; warmup" % t_symbolic)
    for _ in range(a.warmup):
        step(32, a.tol, 30)
    if a.save_tuning:
        with open(a.save_tuning, "w") as f:
            json.dump({"workload": "%dx%dx%d, %d RHS, complex64 factor" % (nx, ny, nz, nrhs),
                       "gemm_speedup": (tune or {}).get("gemm_speedup"),
                       "factor": s.get_tiles(32, 0).tolist(), "solve": s.get_tiles(32, 1, nrhs).tolist()}, f)
    progress("timed steps")
    L.wt_dev_sync()
    s.profile(True)
    L.wt_dev_sync()
    t0 = time.perf_counter()
    res32, it32 = [], []
    for _ in range(a.steps):
        r, it = step(32, a.tol, 30)
        res32.append(r)
        it32.append(it)
    L.wt_dev_sync()
    t32 = (time.perf_counter() - t0) / a.steps
    prof = s.profile_read()
    s.profile(False)
    X32 = np.zeros((n, nrhs), np.complex128)
    L.wt_d2h(X32.ctypes.data_as(ctypes.c_void_p), dX, X32.nbytes)
    # factor-amortised rate: the refined block solve alone, against the
    # factor the last timed step left resident (RHS and solutions in HBM)
    progress("factor-amortised solves")
    L.wt_dev_sync()
    t0 = time.perf_counter()
    for _ in range(3):
        W._lib.check(L.wt_solver_solve(s._h, 32, nrhs, dB, dX, 1, a.tol, 30, ctypes.byref(st)), "solve")
    L.wt_dev_sync()
    t_am = (time.perf_counter() - t0) / 3
    amortised = {"rhs_per_s": nrhs / t_am, "s_per_block": t_am, "refinement_solves": st.iterations,
                 "rel_residual": st.rel_residual}
    # the same step with the right-hand sides and solutions in host memory
    # (the boundary's host-buffer mode): the PCIe-inclusive rate, not `value`
    progress("host-buffer step")
    Xh = np.zeros_like(B)
    L.wt_dev_sync()
    t0 = time.perf_counter()
    W._lib.check(L.wt_solver_assemble(s._h, ctypes.byref(g)), "assemble")
    W._lib.check(L.wt_solver_factor(s._h, 32), "factor")
    W._lib.check(L.wt_solver_solve(s._h, 32, nrhs, B.ctypes.data_as(ctypes.c_void_p),
                                   Xh.ctypes.data_as(ctypes.c_void_p), 0, a.tol, 30, ctypes.byref(st)), "solve")
    L.wt_dev_sync()
    t_host = time.perf_counter() - t0
    del Xh
    s.release(32)

    fp64 = None
    progress("FP32 path %.3f s/step" % t32)
    if not a.no_fp64:
        need = s.bytes_needed(64, nrhs)
        try:
            step(64, 0.0, 1)  # allocation + launch plans (untimed)
            if not a.no_autotune:
                progress("autotuning the complex128 factorization's GEMM tile shapes")
                s.autotune(64)  # one factorization per tile variant (untimed)
            L.wt_dev_sync()
            t0 = time.perf_counter()
            r64, _ = step(64, 0.0, 1)
            L.wt_dev_sync()
            t64 = time.perf_counter() - t0
            # per-kernel times of one more FP64 step (HIP events on the solver stream)
            s.profile(True)
            step(64, 0.0, 1)
            L.wt_dev_sync()
            prof64 = s.profile_read()
            s.profile(False)
            X64 = np.zeros((n, nrhs), np.complex128)
            L.wt_d2h(X64.ctypes.data_as(ctypes.c_void_p), dX, X64.nbytes)
            werr = float(np.max(np.linalg.norm(X32 - X64, axis=0) / np.linalg.norm(X64, axis=0)))
            g64 = {k: prof64[k] for k in ("gemm_factor", "gemm_factor_narrow")}
            g64_ms = sum(v["ms"] for v in g64.values())
            g64_tf = sum(v["flops"] for v in g64.values()) / (g64_ms * 1e-3) / 1e12 if g64_ms > 0 else 0.0
            fp64 = {"rhs_per_s": nrhs / t64, "s_per_step": t64, "rel_residual": r64,
                    "fp32ir_vs_fp64_rel_wavefield_err": werr, "hbm_bytes_estimate": need,
                    "kernel_ms_per_step": {k: v["ms"] for k, v in prof64.items()},
                    "factor_gemm": {"ms_per_step": g64_ms, "achieved": g64_tf, "peak": PEAK_FP64_TFLOPS,
                                    "unit": "TFLOP/s", "frac": g64_tf / PEAK_FP64_TFLOPS,
                                    "note": "8 M N K per complex product over every factorization GEMM of the "
                                            "complex128 step (3M executes 3/4 of it on the FP64 matrix cores)"}}
            s.release(64)
        except RuntimeError as e:
            fp64 = {"error": str(e), "hbm_bytes_estimate": need}

    g_spl = prof["gemm_factor_split"]
    split_dominant = g_spl["ms"] > prof["gemm_factor"]["ms"]
    g_fac = g_spl if split_dominant else prof["gemm_factor"]
    kname = SPLIT_KERNEL if split_dominant else GEMM_KERNEL
    achieved = g_fac["flops"] / (g_fac["ms"] * 1e-3) / 1e12 if g_fac["ms"] > 0 else 0.0
    # every factorization GEMM, whichever tile variant the autotuner gave it
    g_nar = prof["gemm_factor_narrow"]
    g_parts = [prof["gemm_factor"], g_nar, g_spl]
    g_all_ms = sum(g["ms"] for g in g_parts)
    g_all_tf = sum(g["flops"] for g in g_parts) / (g_all_ms * 1e-3) / 1e12 if g_all_ms > 0 else 0.0
    all_gemm = {"achieved": g_all_tf, "frac": g_all_tf / PEAK_FP32_TFLOPS, "ms_per_step": g_all_ms / a.steps,
                "launches_per_step": sum(g["launches"] for g in g_parts) / a.steps,
                "note": "8 M N K per complex product (fp32 complex-product rate) against the FP32 matrix peak"}
    traffic = gemm_traffic(kname)
    # block solve (all right-hand sides at once): factor entries read once per
    # solve, 8 flops per entry and right-hand side; time = the solve GEMMs,
    # the super-panel triangular solves and the gathers / permutations
    n_solves = max(sum(it32), 1)
    sv_ms = (prof["gemm_solve"]["ms"] + prof["panel_trsm"]["ms"] + prof["reorder"]["ms"]) / n_solves
    f_bytes = info["factor_entries"] * 8.0
    sv_flops = 8.0 * nrhs * info["factor_entries"]
    sv_gbs = f_bytes / (sv_ms * 1e-3) / 1e9 if sv_ms > 0 else 0.0
    sv_tf = sv_flops / (sv_ms * 1e-3) / 1e12 if sv_ms > 0 else 0.0
    t_roof = max(f_bytes / (PEAK_HBM_GBS * 1e9), sv_flops / (PEAK_FP32_TFLOPS * 1e12))
    block_solve = {"ms_per_solve": sv_ms, "factor_bytes": f_bytes, "gbs": sv_gbs, "frac_hbm": sv_gbs / PEAK_HBM_GBS,
                   "tflops": sv_tf, "flop_per_byte": sv_flops / f_bytes,
                   "frac_min_roof": (t_roof * 1e3) / sv_ms if sv_ms > 0 else 0.0,
                   "solves_per_step": n_solves / a.steps}
    asm = prof["assemble"]
    asm_gbs = asm["bytes"] / (asm["ms"] * 1e-3) / 1e9 if asm["ms"] > 0 else 0.0
    progress("CPU baseline")
    cpu = None if a.no_cpu else cpu_baseline(nx, ny, nz, nrhs, a.cpu_budget, max(int(round(n_solves / a.steps)), 1))
    if cpu is not None:
        progress("CPU reference assembly")
        cpu["assembly_reference"] = reference_assembly(nx, ny, nz, dx, freq, vp, vs, rho,
                                                       asm["ms"] / max(asm["launches"], 1))
    progress("serving sample")
    serve = None if a.no_serve else serving_sample(a.serve_rate, a.serve_seconds, a.burst_seconds)
    value = nrhs / t32
    out = {
        "metric": METRIC, "value": value, "unit": "RHS/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": t32 * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 factor%s / f64 refinement" % (" (large GEMMs: fp32 products as bf16 splits, fp32 accumulate)"
                                                    if split_dominant else ""), "data": "synthetic",
        "config": {"workload": "heterogeneous %dx%dx%d dx=200m, 0.3 Hz, %d plane-wave RHS, FP32 factor + "
                               "FP64 IR to 1e-8" % (nx, ny, nz, nrhs),
                   "nx": nx, "ny": ny, "nz": nz, "nrhs": nrhs, "unknowns": info["n"], "nnz": info["nnz"],
                   "fronts": info["fronts"], "factor_flops": info["flops"], "parallelism": "1 GPU"},
        "roofline": gemm_roofline(split_dominant, kname, achieved, traffic, g_fac, all_gemm, a.steps),
        "cpu_baseline": cpu,
        "rel_residual": max(res32), "refinement_solves": max(it32),
        "fp64": fp64,
        "assembly": {"gbs": asm_gbs, "frac_hbm": asm_gbs / PEAK_HBM_GBS, "ms": asm["ms"] / max(asm["launches"], 1)},
        "kernel_ms_per_step": {k: v["ms"] / a.steps for k, v in prof.items()},
        "block_solve": block_solve,
        "factor_amortised": amortised,
        "symbolic_s": t_symbolic,
        "autotune": tune,
        "p50_latency_ms": serve["p50_latency_ms"] if serve else None,
        "p99_latency_ms": serve["p99_latency_ms"] if serve else None,
        "serving": serve,
    }
    if fp64 and "rhs_per_s" in fp64:
        out["fp32ir_over_fp64"] = value / fp64["rhs_per_s"]
    out["pcie_inclusive_rhs_per_s"] = nrhs / t_host
    L.wt_dev_free(dB)
    L.wt_dev_free(dX)
    s.close()
    if not a.no_sweep:
        progress("multi-frequency layered sweep")
        out["sweep"] = sweep_sample()
    if not a.no_batch:
        progress("batched small models")
        out["batched"] = batch_sample()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
