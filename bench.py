"""Benchmarks of the MI355X Tensor-BLAS on one GPU; one JSON line on stdout.

Default workload (the headline): recursive TRSM X*L^T = B at n = m = 16384,
nb = 256.  The FP32-accurate split tier (FP16 hi/lo split on MFMA) is timed against
the plain FP32 vector-FMA kernel of the same recursion on the same GPU.  The line
also carries the plain FP16 tier, the split tier with FP64-residual refinement, the
max-relative and normwise error of each tier against an FP64 host solve of a sample of
rows, the dominant kernel's roofline and a host BLAS baseline on a bounded sample.

One step = one tb_trsm call on a copy of B of its own (the solve overwrites B): the
copies are made before the timed region, so inputs are resident in HBM when it starts
(up to 32 copies; steps beyond that restore a copy inside the timed region).

After the headline the same process measures the benchmark's other configurations and
carries each, with its own errors, roofline and host baseline, under "configs"
(--no-configs skips them): SYRK n = k = 2048 (plain / split / FP32-FMA); split SYRK
against FP32-FMA at n = k = 16384; Ozaki SYR2K 32768 x 8192 at phi 0.5 and 2 and Ozaki
SYMM 32768 x 8192 at phi 0.5 against native FP64 MFMA; the 2000-instance batch plan.
Then the request stream (--no-serve skips it) under "served".

Other workloads (--workload):
  syrk   SYRK C = A A^T, lower, n = k = 2048 (--n, --k), nb = 256, plain / split /
         FP32-FMA tiers, error vs FP64 BLAS, host OpenBLAS ssyrk baseline
  ozaki  FP64 SYR2K (--op syr2k) or SYMM (--op symm) by Ozaki INT8 slicing, n = 32768,
         k = 8192, against the native FP64-MFMA tier; error vs FP64 BLAS products of
         64 sampled output columns
  batched  2000 independent SYRK / TRSM instances (n log-uniform in [64, 2048]) as
         one batch plan (persistent kernel, one HIP graph), against the same instances
         as single calls in the same tier and in the FP32-FMA tier
  serve  a stream of mixed requests (SYRK 40 / TRSM 30 / TRMM 15 / SYMM 10 / SYR2K 5 %,
         n log-uniform [128, 8192], tiers plain / split / ozaki) with Poisson arrivals at
         --load of the offline capacity, through the request loop on --slots streams:
         p50/p99 latency against isolated calls, throughput against offline, results
         checked equal to the isolated calls
  potrf  recursive Cholesky n = 16384 (split tier against FP32-FMA and plain), error
         against host FP64 LAPACK dpotrf, host OpenBLAS spotrf baseline
  trtri  triangular inverse n = 16384 (split against FP32-FMA and plain), error against
         host FP64 LAPACK dtrtri, host OpenBLAS strtri baseline
  gemm   tb_gemm microbenchmark (--m --n --k --tiers)
  copy   a 256 MiB device copy (reference for HBM byte counters)

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload trsm]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tensor-blas_amd", "python"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP16_PEAK_TF = 2516.6      # dense FP16 MFMA, MI355X
INT8_PEAK_TOPS = 5033.2    # dense INT8 MFMA
FP64_PEAK_TF = 78.6        # FP64 MFMA
FP32_VEC_PEAK_TF = 157.3
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
# FP16 MFMA products the split tier executes per algorithmic multiply-add, by planes
SPLIT_PRODUCTS = {2: 3, 3: 6}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads():
    try:
        from threadpoolctl import threadpool_info
        return max([i.get("num_threads", 1) for i in threadpool_info()
                    if i.get("internal_api") == "openblas"] or [os.cpu_count() or 1])
    except Exception:
        return os.cpu_count() or 1


def pmc_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` in `workload` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-kernel scaled), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                p = line.split()
                if len(p) >= 4 and not line.startswith("#") and p[0] == kernel and p[1] == workload:
                    return int(float(p[2]) + float(p[3]))
    except OSError:
        pass
    return None


def roofline(prof, peak, unit, bound, workload, per_flop_note):
    """Roofline of the profiled category with the largest summed duration."""
    name, (launches, ms, work) = max(prof.items(), key=lambda kv: kv[1][1])
    achieved = work / (ms * 1e-3) / 1e12
    roof = {"bound": bound, "achieved": round(achieved, 2), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 4), "traffic": pmc_traffic(name, workload)}
    info = {"name": name, "launches": launches, "avg_ms": round(ms / launches, 4),
            "note": "achieved = algorithmic flops of the kernel's launches / their summed "
                    "duration (HIP events on the library stream); " + per_flop_note}
    return roof, info


def timed(h, fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    h.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    h.synchronize()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def profiled(h, fn):
    h.set_profiling(True)
    fn()
    prof = h.profile()
    h.set_profiling(False)
    return prof


def err_metrics(X, R):
    d = np.abs(X.astype(np.float64) - R)
    return {"maxrel": float(d.max() / np.abs(R).max()),
            "normwise": float(np.linalg.norm(d) / np.linalg.norm(R))}


def line_base(metric, value, unit, steps, warmup, ms, dtype, data, config, hib=True):
    return {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1,
            "steps": steps, "warmup": warmup, "ms_per_step": round(ms, 4),
            "higher_is_better": hib, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": data, "config": config}


# ----------------------------------------------------------------------- TRSM
def trsm_inputs(n, m, seed_l=0xC0F2, seed_b=0xC0F3):
    gl = np.random.default_rng(seed_l)
    L = gl.uniform(-1.0, 1.0, (n, n)).astype(np.float32)
    L[np.diag_indices(n)] = (n + gl.uniform(0.0, 1.0, n)).astype(np.float32)
    gb = np.random.default_rng(seed_b)
    B = gb.uniform(-1.0, 1.0, (m, n)).astype(np.float32)
    return L, B     # L's strict upper part is never read (lower storage)


def trsm_host_rows(L, B, rows):
    """FP64 solve of x*L^T = b for the sampled rows (OpenBLAS dtrsm)."""
    import scipy.linalg.blas as sb
    Lf = np.asfortranarray(np.tril(L).astype(np.float64))
    Bs = np.asfortranarray(B[rows].astype(np.float64))
    return sb.dtrsm(1.0, Lf, Bs, side=1, lower=1, trans_a=1)


def trsm_cpu_baseline(L, B, budget_s=20.0):
    """Host OpenBLAS strsm (scipy) on the first rows of the same problem."""
    import scipy.linalg.blas as sb
    n = L.shape[0]
    Lf = np.asfortranarray(L)
    rows = 256
    done_rows, t_total = 0, 0.0
    while t_total < budget_s and done_rows < B.shape[0]:
        Bs = np.asfortranarray(B[done_rows:done_rows + rows])
        t0 = time.perf_counter()
        sb.strsm(1.0, Lf, Bs, side=1, lower=1, trans_a=1, overwrite_b=1)
        t_total += time.perf_counter() - t0
        done_rows += Bs.shape[0]
        if t_total < 2.0:
            rows *= 2
    tf = done_rows * n * n / t_total / 1e12
    return {"value": round(tf, 4), "unit": "TFLOP/s", "cores": int(host_threads()), "kind": "port",
            "impl": "host OpenBLAS strsm via scipy.linalg.blas (right/lower/T/non-unit, FP32)",
            "sample": f"{done_rows} of {B.shape[0]} rows of B, n={n} ({t_total:.1f} s)"}


def run_trsm(args, torch, tensorblas):
    n, m, nb = args.n or 16384, args.m or 16384, args.nb
    log(f"bench: trsm inputs n={n} m={m} nb={nb}")
    L, B = trsm_inputs(n, m)
    dL = torch.from_numpy(np.ascontiguousarray(L.T)).cuda()     # column-major L, lda = n
    dB0 = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()    # column-major B, ldb = m
    dB = torch.empty_like(dB0)
    h = tensorblas.Handle()
    h.set_split_planes(args.split_planes)
    flops = float(m) * n * n

    def step(tier):
        def f():
            dB.copy_(dB0)
            h.trsm(tier, m, n, dL.data_ptr(), n, dB.data_ptr(), m, nb=nb)
        return f

    # The solve overwrites B, so every timed step gets a copy of B of its own, made
    # before the timed region (up to 32 of them, 1 GiB each at 16384^2; past that the
    # steps reuse them and restore B inside the timed region).
    stage = [torch.empty_like(dB0) for _ in range(min(max(args.steps, 3), 32))]

    def timed_staged(tier, steps, warmup):
        for _ in range(warmup):
            step(tier)()
        for b in stage[:steps]:
            b.copy_(dB0)
        h.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            b = stage[i % len(stage)]
            if i >= len(stage):
                b.copy_(dB0)
            h.trsm(tier, m, n, dL.data_ptr(), n, b.data_ptr(), m, nb=nb)
        h.synchronize()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3

    log("bench: split tier (timed)")
    ms_split = timed_staged("split", args.steps, args.warmup)
    rows = np.random.default_rng(7).choice(m, 48, replace=False)
    Xs = stage[0].cpu().numpy().T[rows]
    log("bench: split tier (profiled step)")
    prof = profiled(h, step("split"))
    log("bench: fp32 FMA tier")
    ms_fp32 = timed_staged("fp32", 2, 1)
    Xf = stage[0].cpu().numpy().T[rows]
    log("bench: plain FP16 tier")
    ms_plain = timed_staged("plain", 3, 1)
    Xp = stage[0].cpu().numpy().T[rows]
    log("bench: split tier + FP64-residual refinement")

    def refined():
        dB.copy_(dB0)
        h.trsm_refined("split", m, n, dL.data_ptr(), n, dB.data_ptr(), m, nb=nb)
    ms_ref = timed(h, refined, 3, 1)
    Xr = dB.cpu().numpy().T[rows]
    log("bench: FP64 host reference rows")
    R = trsm_host_rows(L, B, rows)
    errs = {"split": err_metrics(Xs, R), "fp32": err_metrics(Xf, R), "plain": err_metrics(Xp, R),
            "split_refined": err_metrics(Xr, R)}

    gemm_prof = {k: v for k, v in prof.items() if k.startswith("gemm_")}
    products = SPLIT_PRODUCTS[args.split_planes]
    roof, rk = roofline(gemm_prof, FP16_PEAK_TF, "TFLOP/s", "mfma", f"trsm{n}",
                        f"the split tier ({args.split_planes} FP16 planes) executes {products} "
                        f"FP16 MFMA products per algorithmic multiply-add, so its MFMA ceiling in "
                        f"algorithmic flops is {FP16_PEAK_TF / products:.1f} TFLOP/s")
    rk["mfma_executed_frac"] = round(roof["frac"] * products, 4)
    if roof["traffic"]:
        rk["traffic_source"] = ("profiles/hbm_traffic.txt: rocprofv3 --pmc FETCH_SIZE and "
                                "WRITE_SIZE, scaled by a 256 MiB copy")
    value = flops / (ms_split * 1e-3) / 1e12
    line = line_base("eff. TFLOP/s at FP32-level error vs plain, same GPU; served p50/p99 latency",
                     value, "TFLOP/s", args.steps, args.warmup, ms_split,
                     f"f32 ({args.split_planes}-plane fp16 split MFMA)",
                     "synthetic: L strictly-lower U[-1,1], diag n+U[0,1]; B U[-1,1] (seeded)",
                     {"workload": "recursive TRSM X*L^T=B, right/lower/T/non-unit, FP32-accurate "
                                  "split tier", "n": n, "m": m, "nb": nb,
                      "split_planes": args.split_planes})
    line.update({
        "vs_plain_fp32_fma": {"fp32_fma_tflops": round(flops / (ms_fp32 * 1e-3) / 1e12, 2),
                              "ratio": round(ms_fp32 / ms_split, 3)},
        "plain_fp16_tflops": round(flops / (ms_plain * 1e-3) / 1e12, 2),
        "split_refined": {"tflops": round(flops / (ms_ref * 1e-3) / 1e12, 2), "ms": round(ms_ref, 3),
                          "note": "tb_trsm_refined: split solve, FP64-MFMA residual B - X L^T, "
                                  "split solve of the correction (flops counted once, m n^2)"},
        "errors_vs_fp64": errs,
        "roofline": roof,
        "roofline_kernel": rk,
        "profile_ms": {k: round(v[1], 3) for k, v in prof.items()},
    })
    if not args.no_cpu:
        log("bench: host OpenBLAS strsm sample")
        line["cpu_baseline"] = trsm_cpu_baseline(L, B, budget_s=10.0)
    h.close()
    del dL, dB0, dB
    torch.cuda.empty_cache()
    return line


# ----------------------------------------------------------------------- SYRK
def run_syrk(args, torch, tensorblas):
    """SYRK C = A A^T (lower, alpha 1, beta 0) in the plain, split and FP32-FMA tiers.
    n <= 4096: errors over the whole lower triangle against FP64 dsyrk; larger n: over
    32 sampled columns (lower part) against FP64 products.  value = the `--tier` tier
    (default plain at n <= 4096, split above); the roofline is that tier's GEMM."""
    n, k, nb = args.n or 2048, args.k or (args.n or 2048), args.nb
    main = args.tier or ("plain" if n <= 4096 else "split")
    log(f"bench: syrk n={n} k={k}")
    A = np.random.default_rng(0xC0F1).uniform(-1.0, 1.0, (n, k)).astype(np.float32)
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()     # column-major n x k
    dC = torch.zeros(n * n, device="cuda", dtype=torch.float32)
    h = tensorblas.Handle()
    h.set_split_planes(args.split_planes)
    flops = float(n) * n * k

    def step(tier):
        return lambda: h.syrk(tier, n, k, 1.0, dA.data_ptr(), n, 0.0, dC.data_ptr(), n, nb=nb)

    import scipy.linalg.blas as sb
    if n <= 4096:
        R = sb.dsyrk(1.0, np.asfortranarray(A.astype(np.float64)), lower=1)
        low = np.tril_indices(n)

        def err():
            return err_metrics(dC.cpu().numpy().reshape(n, n).T[low], R[low])
        ref_note = "FP64 OpenBLAS dsyrk, whole lower triangle"
    else:
        cols = np.sort(np.random.default_rng(13).choice(n, 32, replace=False))
        A64 = A.astype(np.float64)
        Rc = A64 @ np.ascontiguousarray(A64[cols].T)             # n x 32
        del A64

        def err():
            Cv = dC.view(n, n)
            got = np.concatenate([Cv[j, j:].cpu().numpy() for j in cols])
            ref = np.concatenate([Rc[j:, c] for c, j in enumerate(cols)])
            return err_metrics(got, ref)
        ref_note = "FP64 products of 32 sampled columns (lower part)"
    out, errs = {}, {}
    for tier in ("plain", "split", "fp32"):
        out[tier] = timed(h, step(tier), args.steps if tier == main else 3, args.warmup)
        errs[tier] = err()
    prof = profiled(h, step(main))
    products = SPLIT_PRODUCTS[args.split_planes] if main == "split" else 1
    roof, rk = roofline({k_: v for k_, v in prof.items() if k_.startswith("gemm_")}, FP16_PEAK_TF,
                        "TFLOP/s", "mfma", f"syrk{n}",
                        f"{main} tier, {products} FP16 MFMA product(s) per algorithmic flop"
                        + ("; one launch of a few microseconds: launch and latency bound" if n <= 2048 else ""))
    if products > 1:
        rk["mfma_executed_frac"] = round(roof["frac"] * products, 4)
    ms = out[main]
    line = line_base(f"eff. TFLOP/s, {main} SYRK", flops / (ms * 1e-3) / 1e12, "TFLOP/s",
                     args.steps, args.warmup, ms,
                     "f16 in, f32 accumulate" if main == "plain" else
                     f"f32 ({args.split_planes}-plane fp16 split MFMA)",
                     "synthetic: A U[-1,1] FP32 (seeded)",
                     {"workload": "SYRK C=A*A^T lower, alpha=1, beta=0", "n": n, "k": k, "nb": nb,
                      "tier": main})
    line.update({
        "tiers_tflops": {t: round(flops / (v * 1e-3) / 1e12, 2) for t, v in out.items()},
        "split_vs_fp32_fma": round(out["fp32"] / out["split"], 3),
        "errors_vs_fp64": errs, "reference": ref_note, "roofline": roof, "roofline_kernel": rk,
        "profile_ms": {k_: round(v[1], 4) for k_, v in prof.items()},
    })
    if not args.no_cpu:
        ns = min(n, 8192)
        Af = np.asfortranarray(A[:ns])
        sb.ssyrk(1.0, Af[:256], lower=1)
        ts = []
        for _ in range(3 if ns <= 4096 else 1):
            t0 = time.perf_counter()
            sb.ssyrk(1.0, Af, lower=1)
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        line["cpu_baseline"] = {"value": round(float(ns) * ns * k / t / 1e12, 4), "unit": "TFLOP/s",
                                "cores": int(host_threads()), "kind": "port",
                                "impl": "host OpenBLAS ssyrk via scipy.linalg.blas",
                                "sample": (f"full problem, median of 3 ({t * 1e3:.1f} ms)" if ns == n else
                                           f"first {ns} rows of A (n={ns}, k={k}), one run ({t:.2f} s)")}
    h.close()
    del dA, dC
    torch.cuda.empty_cache()
    return line


# ------------------------------------------------------------------- Cholesky
def run_potrf(args, torch, tensorblas):
    """Recursive Cholesky A = L L^T (lower) of a symmetric diagonally dominant matrix
    (U[-1,1] symmetric, diagonal + n: positive definite), split tier against the
    FP32-FMA and plain tiers; error of L against the host FP64 LAPACK factor."""
    n, nb = args.n or 16384, args.nb
    g = np.random.default_rng(0xC0F8)
    A = g.uniform(-1.0, 1.0, (n, n)).astype(np.float32)
    A = np.tril(A) + np.tril(A, -1).T
    A[np.diag_indices(n)] += np.float32(n)
    dA0 = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()     # column-major, symmetric
    dA = dA0.clone()
    h = tensorblas.Handle()
    h.set_split_planes(args.split_planes)
    flops = float(n) ** 3 / 3.0
    infos = []

    def step(tier):
        def f():
            dA.copy_(dA0)
            infos.append(h.potrf(tier, n, dA.data_ptr(), n, nb=nb))
        return f
    log("bench: potrf host FP64 reference")
    import scipy.linalg.lapack as sl
    R, rinfo = sl.dpotrf(np.asfortranarray(A.astype(np.float64)), lower=1)
    R = np.tril(R)
    out, errs = {}, {}
    for tier in ("split", "fp32", "plain"):
        out[tier] = timed(h, step(tier), args.steps if tier == "split" else 3, args.warmup)
        L = np.tril(dA.cpu().numpy().reshape(n, n).T).astype(np.float64)
        d = np.abs(L - R)
        errs[tier] = {"normwise": float(np.linalg.norm(d) / np.linalg.norm(R)),
                      "maxrel": float(d.max() / np.abs(R).max())}
    assert rinfo == 0 and all(i == 0 for i in infos), (rinfo, set(infos))
    prof = profiled(h, step("split"))
    roof, rk = roofline({k_: v for k_, v in prof.items() if k_.startswith("gemm_")}, FP16_PEAK_TF,
                        "TFLOP/s", "mfma", f"potrf{n}",
                        "split tier (2 FP16 planes): 3 FP16 MFMA products per algorithmic flop")
    ms = out["split"]
    line = line_base("eff. TFLOP/s of recursive Cholesky at FP32-level error (n^3/3 flops)",
                     flops / (ms * 1e-3) / 1e12, "TFLOP/s", args.steps, args.warmup, ms,
                     f"f32 ({args.split_planes}-plane fp16 split MFMA)",
                     "synthetic: A symmetric U[-1,1], diagonal + n (seeded)",
                     {"workload": "POTRF A = L L^T lower, recursive (TRSM + SYRK + diagonal blocks in FP32)",
                      "n": n, "nb": nb})
    line.update({
        "tiers_tflops": {t: round(flops / (v * 1e-3) / 1e12, 2) for t, v in out.items()},
        "split_vs_fp32_fma": round(out["fp32"] / out["split"], 3),
        "errors_vs_fp64": errs, "roofline": roof, "roofline_kernel": rk,
        "profile_ms": {k_: round(v[1], 3) for k_, v in prof.items()},
        "note": "each step restores A (device copy) before the factorization",
    })
    if not args.no_cpu:
        Af = np.asfortranarray(A)
        t0 = time.perf_counter()
        sl.spotrf(Af, lower=1)
        t = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(flops / t / 1e12, 4), "unit": "TFLOP/s",
                                "cores": int(host_threads()), "kind": "port",
                                "impl": "host OpenBLAS spotrf via scipy.linalg.lapack",
                                "sample": f"full problem, one run ({t:.2f} s)"}
    h.close()
    return line


# --------------------------------------------------------- triangular inverse
def run_trtri(args, torch, tensorblas):
    """In-place inverse of the TRSM benchmark's lower-triangular L (strictly lower
    U[-1,1], diagonal n + U[0,1]), split tier against the FP32-FMA and plain tiers;
    error against the host FP64 LAPACK inverse."""
    n, nb = args.n or 16384, args.nb
    g = np.random.default_rng(0xC0F9)
    L = (np.tril(g.uniform(-1.0, 1.0, (n, n)), -1) + np.diag(n + g.uniform(0, 1, n))).astype(np.float32)
    dA0 = torch.from_numpy(np.ascontiguousarray(L.T)).cuda()     # column-major
    dA = dA0.clone()
    h = tensorblas.Handle()
    h.set_split_planes(args.split_planes)
    flops = float(n) ** 3 / 3.0
    infos = []

    def step(tier):
        def f():
            dA.copy_(dA0)
            infos.append(h.trtri(tier, "L", "N", n, dA.data_ptr(), n, nb=nb))
        return f
    log("bench: trtri host FP64 reference")
    import scipy.linalg.lapack as sl
    R, rinfo = sl.dtrtri(np.asfortranarray(L.astype(np.float64)), lower=1)
    R = np.tril(R)
    out, errs = {}, {}
    for tier in ("split", "fp32", "plain"):
        out[tier] = timed(h, step(tier), args.steps if tier == "split" else 3, args.warmup)
        X = np.tril(dA.cpu().numpy().reshape(n, n).T).astype(np.float64)
        d = np.abs(X - R)
        errs[tier] = {"normwise": float(np.linalg.norm(d) / np.linalg.norm(R)),
                      "maxrel": float(d.max() / np.abs(R).max())}
    assert rinfo == 0 and all(i == 0 for i in infos), (rinfo, set(infos))
    prof = profiled(h, step("split"))
    roof, rk = roofline({k_: v for k_, v in prof.items() if k_.startswith("gemm_")}, FP16_PEAK_TF,
                        "TFLOP/s", "mfma", f"trtri{n}",
                        "split tier (2 FP16 planes): 3 FP16 MFMA products per algorithmic flop")
    ms = out["split"]
    line = line_base("eff. TFLOP/s of the triangular inverse at FP32-level error (n^3/3 flops)",
                     flops / (ms * 1e-3) / 1e12, "TFLOP/s", args.steps, args.warmup, ms,
                     f"f32 ({args.split_planes}-plane fp16 split MFMA)",
                     "synthetic: L strictly lower U[-1,1], diagonal n + U[0,1] (seeded)",
                     {"workload": "TRTRI inv(L) lower in place, recursive (TRMM + diagonal blocks in FP32)",
                      "n": n, "nb": nb})
    line.update({
        "tiers_tflops": {t: round(flops / (v * 1e-3) / 1e12, 2) for t, v in out.items()},
        "split_vs_fp32_fma": round(out["fp32"] / out["split"], 3),
        "errors_vs_fp64": errs, "roofline": roof, "roofline_kernel": rk,
        "profile_ms": {k_: round(v[1], 3) for k_, v in prof.items()},
        "note": "each step restores L (device copy) before the inversion",
    })
    if not args.no_cpu:
        Lf = np.asfortranarray(L)
        t0 = time.perf_counter()
        sl.strtri(Lf, lower=1)
        t = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(flops / t / 1e12, 4), "unit": "TFLOP/s",
                                "cores": int(host_threads()), "kind": "port",
                                "impl": "host OpenBLAS strtri via scipy.linalg.lapack",
                                "sample": f"full problem, one run ({t:.2f} s)"}
    h.close()
    return line


# ---------------------------------------------------------------------- Ozaki
OZAKI_DEFAULT_SLICES = {0.5: "8,8+,9", 2.0: "8+,9+,10"}


def slice_products(S, extra):
    """INT8 MFMA slice products per FP64 multiply-add: the triangular scheme's
    S(S+1)/2, plus the S-1 pairs of the group g = S when extra."""
    return S * (S + 1) // 2 + extra * (S - 1)


def int_slices(X, bits, nslices):
    """X[i, :] = sum_s S[s][i, :] * 2^(e[i] - bits*(s+1)) + (a remainder below
    2^(e[i] - bits*nslices)), each S[s] an integer-valued float64 array with
    |S[s]| < 2^bits, e[i] the exponent with max|X[i, :]| < 2^e[i].  Every step (scaling
    by a power of two, truncation, subtraction of the integer part) is exact."""
    mx = np.abs(X).max(axis=1)
    e = np.where(mx > 0, np.frexp(mx)[1], 0).astype(np.int64)
    R = X * np.ldexp(1.0, bits - e)[:, None]
    out = []
    for _ in range(nslices):
        a = np.trunc(R)
        out.append(a)
        R -= a
        R *= 2.0 ** bits
    return out, e


def exact_rowdot(pairs, cols):
    """out[:, c] = sum over (X, Y) in pairs of X @ Y[cols[c]], to within a few FP64 ulps:
    rows of X and the sampled rows of Y are cut into three integer slices of b bits
    (3b >= 57) with 2b + log2(k) <= 53, so every FP64 BLAS product of two slices sums
    integers below 2^53 exactly; the nine (eighteen) exact slice products are scaled
    and summed in long double and rounded once."""
    k = pairs[0][0].shape[1]
    bits = (53 - int(np.ceil(np.log2(max(k, 2))))) // 2
    acc = None
    for X, Y in pairs:
        xs, ex = int_slices(X, bits, 3)
        ys, ey = int_slices(Y[cols], bits, 3)
        for s_, xa in enumerate(xs):
            for t_, yb in enumerate(ys):
                G = (xa @ yb.T).astype(np.longdouble)                       # exact integers
                sc = np.ldexp(np.longdouble(1.0), (ex[:, None] + ey[None, :] - bits * (s_ + t_ + 2)))
                acc = G * sc if acc is None else acc + G * sc
        del xs, ys
    return acc.astype(np.float64)


def run_ozaki(args, torch, tensorblas):
    """SYR2K C = A B^T + B A^T (n x k operands) or SYMM C = A B (A m x m symmetric,
    lower stored; B m x n), FP64 by Ozaki INT8 slicing at several slice counts and by
    native FP64 MFMA.  Entries U[-0.5,0.5]*exp(phi*N(0,1)), generated on the device from
    a seeded generator.  Errors: max-relative and normwise over 64 sampled output columns
    against exact_rowdot's products of the same columns (exact integer slice products
    on FP64 BLAS, accurate to a few ulps; a plain FP64 product would itself be off by
    about 1e-15 of max|C| here, too close to the 1e-14 target to decide it)."""
    op = args.op or "syr2k"
    n, k = args.n or 32768, args.k or 8192
    phi = args.phi
    gen = torch.Generator(device="cuda")
    gen.manual_seed((0xC0F4 if op == "syr2k" else 0xC0F5) + int(round(phi * 10)))

    def mat(r, c):
        u = torch.rand(r, c, device="cuda", dtype=torch.float64, generator=gen) - 0.5
        return u.mul_(torch.randn(r, c, device="cuda", dtype=torch.float64, generator=gen).mul_(phi).exp_())
    ncols = 64
    cols = np.sort(np.random.default_rng(11).choice(n if op == "syr2k" else k, ncols, replace=False))
    log(f"bench: ozaki {op} n={n} k={k} phi={phi}")
    if op == "syr2k":
        # tensor rows are the columns of the column-major n x k operands
        dA, dB = mat(k, n), mat(k, n)
        dC = torch.zeros(n * n, device="cuda", dtype=torch.float64)
        flops = 2.0 * n * n * k
        A = dA.cpu().numpy().T          # n x k views (F order)
        B = dB.cpu().numpy().T
        # C(:, j) = A B(j,:)^T + B A(j,:)^T, all rows (lower part compared)
        ref = exact_rowdot([(np.ascontiguousarray(A), np.ascontiguousarray(B)),
                            (np.ascontiguousarray(B), np.ascontiguousarray(A))], cols)

        def step(tier):
            return lambda: h.syr2k(tier, n, k, 1.0, dA.data_ptr(), n, dB.data_ptr(), n, 0.0,
                                   dC.data_ptr(), n, nb=256)

        def out_col(c, j):
            return dC.view(n, n)[j, j:].cpu().numpy(), ref[j:, c]
        shape = {"n": n, "k": k}
        sample = 8192

        def host_leg():
            import scipy.linalg.blas as sb
            As, Bs = np.asfortranarray(A[:sample]), np.asfortranarray(B[:sample])
            sb.dsyr2k(1.0, As[:512], Bs[:512], lower=1)
            t0 = time.perf_counter()
            sb.dsyr2k(1.0, As, Bs, lower=1)
            t = time.perf_counter() - t0
            return 2.0 * sample * sample * k / t / 1e12, t, f"dsyr2k n={sample} (first rows), k={k}"
    else:
        m, nn = n, k
        dA = mat(m, m)                    # column-major m x m: only the lower triangle is read
        dB = mat(nn, m)                   # column-major m x nn
        dC = torch.zeros(m * nn, device="cuda", dtype=torch.float64)
        flops = 2.0 * m * m * nn
        Mc = dA.cpu().numpy()             # Mc[j, i] = A(i, j): A(i, j) = Mc[min, max]
        B = dB.cpu().numpy().T            # m x nn
        # rows of the symmetric A: A(i, j) = Mc[min(i,j), max(i,j)]
        Asym = np.triu(Mc)
        Asym += np.triu(Mc, 1).T
        ref = exact_rowdot([(Asym, np.ascontiguousarray(B.T))], cols)
        del Asym

        def step(tier):
            return lambda: h.symm(tier, m, nn, 1.0, dA.data_ptr(), m, dB.data_ptr(), m, 0.0,
                                  dC.data_ptr(), m)

        def out_col(c, j):
            return dC.view(nn, m)[j].cpu().numpy(), ref[:, c]
        shape = {"m": m, "n": nn}
        sample = 8192

        def host_leg():
            import scipy.linalg.blas as sb
            Asub = np.asfortranarray(np.triu(Mc[:sample, :sample]).T)   # lower-stored block
            Bsub = np.asfortranarray(B[:sample, :sample])
            sb.dsymm(1.0, Asub[:512, :512], Bsub[:512, :512], side=0, lower=1)
            t0 = time.perf_counter()
            sb.dsymm(1.0, Asub, Bsub, side=0, lower=1)
            t = time.perf_counter() - t0
            return 2.0 * sample * sample * sample / t / 1e12, t, f"dsymm m={sample} n={sample} (leading block)"
    h = tensorblas.Handle()

    def col_err():
        worst, num, den = 0.0, 0.0, 0.0
        for c, j in enumerate(cols):
            got, r = out_col(c, j)
            d = np.abs(got - r)
            worst = max(worst, float(d.max() / np.abs(r).max()))
            num += float((d ** 2).sum())
            den += float((r ** 2).sum())
        return {"maxrel": worst, "normwise": (num / den) ** 0.5}

    res = {}
    # "8" = 8 slices, triangular scheme; "8+" = 8 slices plus the slice-pair group g = 8
    slices = args.slices or OZAKI_DEFAULT_SLICES.get(phi, "8,8+,9")
    for tok in slices.split(","):
        s_, extra = int(tok.rstrip("+")), int(tok.endswith("+"))
        h.set_ozaki_slices(s_)
        h.set_ozaki_extra(extra)
        ms = timed(h, step("ozaki"), args.steps, args.warmup)
        res[f"ozaki_s{tok}"] = {"ms": round(ms, 2), "tflops": round(flops / (ms * 1e-3) / 1e12, 2),
                                "int8_products": slice_products(s_, extra),
                                "errors_vs_ref": col_err()}
        log(f"bench: ozaki {op} phi={phi} S={tok}: {res[f'ozaki_s{tok}']}")
    ms64 = timed(h, step("fp64"), args.steps, args.warmup)
    res["fp64_mfma"] = {"ms": round(ms64, 2), "tflops": round(flops / (ms64 * 1e-3) / 1e12, 2),
                        "errors_vs_ref": col_err()}
    ok = [(v["ms"], key) for key, v in res.items()
          if key != "fp64_mfma" and v["errors_vs_ref"]["maxrel"] <= 1e-14]
    best = min(ok) if ok else min((v["ms"], key) for key, v in res.items() if key != "fp64_mfma")
    tbest = best[1].split("_s")[1]
    sbest, xbest = int(tbest.rstrip("+")), int(tbest.endswith("+"))
    h.set_ozaki_slices(sbest)
    h.set_ozaki_extra(xbest)
    prof = profiled(h, step("ozaki"))
    products = slice_products(sbest, xbest)
    oz = {k_: (v[0], v[1], v[2] * products) for k_, v in prof.items() if k_ == "ozaki_gemm"}
    roof, rk = roofline(oz, INT8_PEAK_TOPS, "TOPS", "mfma", f"{op}{n}x{k}_s{tbest}",
                        f"INT8 MFMA operations = FP64-equivalent flops x {products} slice products "
                        f"({sbest} slices{' + group g=S' if xbest else ''}); the FP64-equivalent "
                        f"ceiling is {INT8_PEAK_TOPS / products:.1f} TFLOP/s")
    rk["fp64_equiv_tflops"] = round(roof["achieved"] / products, 2)
    ms = best[0]
    line = line_base(f"eff. FP64 TFLOP/s of Ozaki {op.upper()} at <=1e-14 max relative error",
                     flops / (ms * 1e-3) / 1e12, "TFLOP/s", args.steps, args.warmup, ms,
                     "f64 (int8 slices, fp64 accumulate)",
                     f"synthetic: U[-0.5,0.5]*exp({phi}*N(0,1)) FP64 (seeded, generated on the device)",
                     {"workload": f"{op.upper()}, Ozaki INT8 slicing", **shape, "phi": phi,
                      "slices": sbest, "extra_group": xbest})
    e64 = res["fp64_mfma"]["errors_vs_ref"]["maxrel"]
    line.update({"results": res, "vs_fp64_mfma": round(ms64 / ms, 3), "meets_1e-14": bool(ok),
                 "fp64_mfma_meets_1e-14": bool(e64 <= 1e-14),
                 "roofline": roof, "roofline_kernel": rk,
                 "profile_ms": {k_: round(v[1], 3) for k_, v in prof.items()},
                 "reference": f"exact integer-slice products of {ncols} sampled output columns"})
    if not args.no_cpu:
        tf, t, what = host_leg()
        line["cpu_baseline"] = {"value": round(tf, 4), "unit": "TFLOP/s", "cores": int(host_threads()),
                                "kind": "port", "impl": "host OpenBLAS via scipy.linalg.blas (FP64)",
                                "sample": f"{what}, one run ({t:.2f} s)"}
    h.close()
    del dA, dB, dC
    torch.cuda.empty_cache()
    return line


# -------------------------------------------------------------------- batched
def batched_instances(count, seed=0xC0F6):
    """count instances: half SYRK (k = n), half TRSM (m = n), n log-uniform in
    [64, 2048] rounded to an integer, shuffled."""
    g = np.random.default_rng(seed)
    ns = np.rint(np.exp(g.uniform(np.log(64), np.log(2048), count))).astype(np.int64)
    ops = np.array(["syrk"] * (count // 2) + ["trsm"] * (count - count // 2))
    g.shuffle(ops)
    return list(zip(ops.tolist(), ns.tolist()))


def run_batched(args, torch, tensorblas):
    count = args.count or 2000
    tier = args.tier or "split"
    inst = batched_instances(count)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0xC0F6)
    nin = sum(n * n for _, n in inst)
    nb_out = sum(n * n for op, n in inst if op == "trsm")
    nc_out = sum(n * n for op, n in inst if op == "syrk")
    log(f"bench: batched {count} instances, {4 * (nin + 2 * nb_out + nc_out) / 2**30:.1f} GiB")
    ins = torch.rand(nin, device="cuda", generator=gen) * 2 - 1        # A (SYRK) / L (TRSM)
    bout = torch.rand(max(1, nb_out), device="cuda", generator=gen) * 2 - 1   # B (TRSM)
    bsave = bout.clone()
    cout = torch.zeros(max(1, nc_out), device="cuda")                # C (SYRK, beta = 0)
    entries, views = [], []
    oi = ob = oc = 0
    flops = 0.0
    for op, n in inst:
        a = ins[oi:oi + n * n]
        oi += n * n
        if op == "trsm":
            L = a.view(n, n)          # column-major L: row j of this view is column j
            L.diagonal().copy_(n + torch.rand(n, device="cuda", generator=gen))
            o = bout[ob:ob + n * n]
            ob += n * n
            entries.append(dict(op="trsm", m=n, n=n, A=a.data_ptr(), lda=n, C=o.data_ptr(), ldc=n))
        else:
            o = cout[oc:oc + n * n]
            oc += n * n
            entries.append(dict(op="syrk", n=n, k=n, A=a.data_ptr(), lda=n, C=o.data_ptr(), ldc=n,
                                alpha=1.0, beta=0.0))
        views.append((op, n, a, o))
        flops += float(n) ** 3
    h = tensorblas.Handle()
    h.set_split_planes(args.split_planes)
    t0 = time.perf_counter()
    plan = tensorblas.Batch(h, tier, entries, nb=args.nb, autotune=True)
    plan_s = time.perf_counter() - t0
    info = plan.info()
    log(f"bench: plan {plan_s:.2f} s, {info}")

    def step():
        bout.copy_(bsave)
        plan.run()
    ms = timed(h, step, args.steps, args.warmup)
    prof = profiled(h, step)
    roof, rk = roofline(prof, FP16_PEAK_TF, "TFLOP/s", "mfma", f"batched{count}",
                        f"one graph run: batched exponent and inverse prologues + the persistent "
                        f"kernel; {tier} tier")

    # errors of sampled instances against FP64 host BLAS (outputs of the last run)
    import scipy.linalg.blas as sb
    errs = []
    pick = [i for i, (op, n) in enumerate(inst) if op == "syrk"][:3] + \
           [i for i, (op, n) in enumerate(inst) if op == "trsm"][:3]
    for i in pick:
        op, n, a, o = views[i]
        A = a.view(n, n).cpu().numpy().T.astype(np.float64)
        got = o.view(n, n).cpu().numpy().T
        if op == "syrk":
            R = sb.dsyrk(1.0, np.asfortranarray(A), lower=1)
            lo = np.tril_indices(n)
            errs.append(dict(op=op, n=n, **err_metrics(got[lo], R[lo])))
        else:
            j0 = sum(m * m for p, m in inst[:i] if p == "trsm")
            B0 = bsave[j0:j0 + n * n].view(n, n).cpu().numpy().T.astype(np.float64)
            R = sb.dtrsm(1.0, np.asfortranarray(np.tril(A)), np.asfortranarray(B0), side=1,
                         lower=1, trans_a=1)
            errs.append(dict(op=op, n=n, **err_metrics(got, R)))

    # the same instances as one single call each (same tier; and the FP32-FMA tier)
    def loop(t):
        def f():
            bout.copy_(bsave)
            for (op, n, a, o), e in zip(views, entries):
                if op == "syrk":
                    h.syrk(t, n, n, 1.0, e["A"], n, 0.0, e["C"], n, nb=args.nb)
                else:
                    h.trsm(t, n, n, e["A"], n, e["C"], n, nb=args.nb)
        return f
    ms_loop = timed(h, loop(tier), 1, 1)
    ms_fp32 = timed(h, loop("fp32"), 1, 1)
    line = line_base("eff. TFLOP/s of a batched mixed SYRK+TRSM plan (one graph launch)",
                     flops / (ms * 1e-3) / 1e12, "TFLOP/s", args.steps, args.warmup, ms,
                     f"f32 ({tier} tier, {args.split_planes}-plane fp16 split MFMA)" if tier == "split"
                     else "f16 in, f32 accumulate",
                     "synthetic: A, B U[-1,1]; L U[-1,1] with diag n+U[0,1] (seeded, on device)",
                     {"workload": f"{count} independent instances, {count // 2} SYRK (k=n) + "
                                  f"{count - count // 2} TRSM (m=n), n log-uniform [64, 2048]",
                      "count": count, "nb": args.nb, "tier": tier, "flops": flops})
    line.update({
        "plan": {"seconds": round(plan_s, 3), **info},
        "vs_single_calls": {"ms": round(ms_loop, 3), "tflops": round(flops / (ms_loop * 1e-3) / 1e12, 2),
                            "speedup": round(ms_loop / ms, 3)},
        "vs_plain_fp32_fma": {"ms": round(ms_fp32, 3),
                              "tflops": round(flops / (ms_fp32 * 1e-3) / 1e12, 2),
                              "ratio": round(ms_fp32 / ms, 3)},
        "errors_vs_fp64": errs, "roofline": roof, "roofline_kernel": rk,
        "note": "each step restores the TRSM right-hand sides (device copy) before the run",
    })
    if not args.no_cpu:
        ts, fl, done = 0.0, 0.0, 0
        for (op, n, a, o), e in zip(views, entries):
            A = np.asfortranarray(a.view(n, n).cpu().numpy().T)
            t0 = time.perf_counter()
            if op == "syrk":
                sb.ssyrk(1.0, A, lower=1)
            else:
                sb.strsm(1.0, A, np.asfortranarray(o.view(n, n).cpu().numpy().T), side=1, lower=1,
                         trans_a=1)
            ts += time.perf_counter() - t0
            fl += float(n) ** 3
            done += 1
            if ts > args.cpu_budget:
                break
        line["cpu_baseline"] = {"value": round(fl / ts / 1e12, 4), "unit": "TFLOP/s",
                                "cores": int(host_threads()), "kind": "port",
                                "impl": "host OpenBLAS ssyrk / strsm via scipy.linalg.blas",
                                "sample": f"first {done} of {count} instances ({ts:.1f} s)"}
    plan.close()
    h.close()
    del ins, bout, bsave, cout, views
    torch.cuda.empty_cache()
    return line


# ---------------------------------------------------------------------- serve
SERVE_MIX = [("syrk", 0.40), ("trsm", 0.30), ("trmm", 0.15), ("symm", 0.10), ("syr2k", 0.05)]
SERVE_FLOPS = {"syrk": 1.0, "trsm": 1.0, "trmm": 1.0, "symm": 2.0, "syr2k": 2.0}   # x n^3


def serve_requests(torch, count, seed=0xC0F7, nmin=128, nmax=8192):
    """count requests: op by the mix, n log-uniform [nmin, nmax] (k = n), tier uniform
    over plain / split / ozaki; operands U[-1,1] on the device (TRSM: diag n+U[0,1])."""
    g = np.random.default_rng(seed)
    ops = g.choice([o for o, _ in SERVE_MIX], size=count, p=[p for _, p in SERVE_MIX])
    ns = np.rint(np.exp(g.uniform(np.log(nmin), np.log(nmax), count))).astype(np.int64)
    tiers = g.choice(["plain", "split", "ozaki"], size=count)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    reqs, keep = [], []
    for op, n, tier in zip(ops.tolist(), ns.tolist(), tiers.tolist()):
        dt = torch.float64 if tier == "ozaki" else torch.float32

        def rnd(n_):
            return torch.rand(n_ * n_, device="cuda", dtype=dt, generator=gen) * 2 - 1
        A, C = rnd(n), rnd(n)
        B = rnd(n) if op in ("syr2k", "trmm", "symm") else None
        r = dict(op=op, tier=tier, m=n, n=n, k=n, nb=256, alpha=1.0, beta=0.0, A=A.data_ptr(),
                 lda=n, B=B.data_ptr() if B is not None else 0, ldb=n, C=C.data_ptr(), ldc=n,
                 flops=SERVE_FLOPS[op] * float(n) ** 3)
        C0 = None
        if op == "trsm":
            A.view(n, n).diagonal().copy_(n + torch.rand(n, device="cuda", dtype=dt, generator=gen))
            C0 = C.clone()                # B is solved in place: restored before each pass
        reqs.append(r)
        keep.append((A, B, C, C0))
    return reqs, keep


def checksum(torch, x):
    v = x.view(torch.int32 if x.dtype == torch.float32 else torch.int64).to(torch.int64)
    w = torch.arange(1, v.numel() + 1, device=v.device, dtype=torch.int64) % 1000003
    return (int(v.sum().item()), int((v * w).sum().item()))


def pct(x, q):
    return float(np.percentile(np.asarray(x), q))


def run_serve(args, torch, tensorblas):
    count = args.count or 500
    nslots = args.slots
    reqs, keep = serve_requests(torch, count)
    total_flops = sum(r["flops"] for r in reqs)
    log(f"bench: serve {count} requests, {total_flops / 1e12:.2f} TFLOP")

    def restore():
        for _, _, C, C0 in keep:
            if C0 is not None:
                C.copy_(C0)
        torch.cuda.synchronize()
    slots = [tensorblas.Handle(torch.cuda.Stream()) for _ in range(nslots)]
    budget = int(args.budget_gib * 2 ** 30)
    # warm-up on slot 0 sizes its workspace; every slot then reserves the same
    for r in reqs:
        tensorblas.serve(slots[:1], budget, [dict(r, arrival_us=0.0)])
    ws = slots[0].workspace_sizes()
    for h in slots[1:]:
        h.reserve_workspace(*ws)
    # isolated: each request alone on an idle GPU
    restore()
    iso = []
    for r in reqs:
        iso.append(tensorblas.serve(slots[:1], budget, [dict(r, arrival_us=0.0)])[0]["done_us"])
    alone = [checksum(torch, C) for _, _, C, _ in keep]
    # offline: everything available at t = 0, all slots, one lane (no admission limit
    # beyond slots and budget): the best throughput the device reaches on this mix
    restore()
    off = tensorblas.serve(slots, budget, [dict(r, arrival_us=0.0) for r in reqs])
    makespan = max(x["done_us"] for x in off)
    cap = count / makespan                                      # requests per microsecond
    # served: Poisson arrivals at args.load of the offline capacity, three admission lanes
    restore()
    g = np.random.default_rng(0xC0F7 + 1)
    arr = np.cumsum(g.exponential(1.0 / (args.load * cap), count)) - 0.0
    arr -= arr[0]
    res = tensorblas.serve(slots, budget, [dict(r, arrival_us=float(a)) for r, a in zip(reqs, arr)],
                           large_us=args.large_ms * 1e3, max_large=args.max_large, order=args.order,
                           reserve_cus=args.reserve_cus, partition=args.partition,
                           small_us=args.small_ms * 1e3, small_slots=args.small_slots)
    lat = [x["done_us"] - a for x, a in zip(res, arr)]
    ratio = [l_ / i_ for l_, i_ in zip(lat, iso)]
    mix_sum = [checksum(torch, C) for _, _, C, _ in keep]
    mismatch = sum(1 for a, b in zip(alone, mix_sum) if a != b)
    span = max(x["done_us"] for x in res) - float(arr[0])
    thr = total_flops / (span * 1e-6) / 1e12
    thr_off = total_flops / (makespan * 1e-6) / 1e12
    by_tier = {}
    for t in ("plain", "split", "ozaki"):
        li = [l_ for l_, r in zip(lat, reqs) if r["tier"] == t]
        by_tier[t] = {"count": len(li), "p50_ms": round(pct(li, 50) / 1e3, 3),
                      "p99_ms": round(pct(li, 99) / 1e3, 3)}
    line = line_base("served throughput TFLOP/s with p50/p99 latency (Poisson arrivals)", thr,
                     "TFLOP/s", 1, 1, span / 1e3, "per request tier: f16 plain / f32 split / f64 Ozaki",
                     "synthetic: U[-1,1] operands resident in HBM (TRSM: diag n+U[0,1]), seeded",
                     {"workload": f"{count} requests, mix SYRK 40/TRSM 30/TRMM 15/SYMM 10/SYR2K 5, "
                                  f"n log-uniform [128, 8192], k=n, tier uniform over plain/split/ozaki",
                      "slots": nslots, "hw_queues": args.hw_queues,
                      "hbm_budget_gib": args.budget_gib, "load": args.load,
                      "large_ms": args.large_ms, "max_large": args.max_large, "order": args.order,
                      "reserve_cus": args.reserve_cus, "partition": args.partition,
                      "small_ms": args.small_ms, "small_slots": args.small_slots,
                      "total_flops": total_flops})
    line.update({
        "latency_ms": {"p50": round(pct(lat, 50) / 1e3, 3), "p99": round(pct(lat, 99) / 1e3, 3)},
        "isolated_latency_ms": {"p50": round(pct(iso, 50) / 1e3, 3), "p99": round(pct(iso, 99) / 1e3, 3)},
        "latency_over_isolated": {"p50": round(pct(ratio, 50), 3), "p99": round(pct(ratio, 99), 3)},
        "p99_latency_vs_isolated_p99": round(pct(lat, 99) / pct(iso, 99), 3),
        "offline": {"makespan_ms": round(makespan / 1e3, 3), "tflops": round(thr_off, 3)},
        "throughput_vs_offline": round(thr / thr_off, 3),
        "by_tier": by_tier,
        "results_equal_alone": {"checked": count, "mismatches": mismatch},
        "refused": sum(1 for x in res if x["status"] == 3),
        "errors": sum(1 for x in res if x["status"] not in (0, 3)),
    })
    if args.dump:
        with open(args.dump, "w") as f:
            json.dump([{"op": r["op"], "tier": r["tier"], "n": r["n"], "flops": r["flops"],
                        "est_us": x["est_us"], "iso_us": i_, "arrival_us": float(a),
                        "admit_us": x["admit_us"], "done_us": x["done_us"], "slot": x["slot"],
                        "offline_done_us": o["done_us"]}
                       for r, x, i_, a, o in zip(reqs, res, iso, arr, off)], f)
    for h in slots:
        h.close()
    return line


# ----------------------------------------------------------------- microbench
def run_gemm(args, torch, tensorblas):
    M, N, K = args.m or 8192, args.n or 8192, args.k or 8192
    h = tensorblas.Handle()
    h.set_split_planes(args.split_planes)
    res = {}
    for tier in args.tiers.split(","):
        dt = torch.float64 if tier in ("fp64", "ozaki") else torch.float32
        A = torch.rand(K, M, device="cuda", dtype=dt) * 2 - 1   # column-major M x K
        B = torch.rand(N, K, device="cuda", dtype=dt) * 2 - 1   # column-major K x N
        C = torch.zeros(N, M, device="cuda", dtype=dt)

        def f():
            h.gemm(tier, 0, 0, M, N, K, 1.0, A.data_ptr(), M, B.data_ptr(), K, 0.0, C.data_ptr(), M)
        ms = timed(h, f, args.steps, args.warmup)
        prof = profiled(h, f)
        res[tier] = {"ms": round(ms, 3), "tflops": round(2.0 * M * N * K / (ms * 1e-3) / 1e12, 2),
                     "kernels": {k_: {"avg_ms": round(v[1] / v[0], 4),
                                      "tflops": round(v[2] / (v[1] * 1e-3) / 1e12, 2)}
                                 for k_, v in prof.items()}}
        log(f"bench: gemm {tier} {M}x{N}x{K}: {res[tier]}")
        del A, B, C
    first = res[args.tiers.split(",")[0]]
    line = line_base("GEMM TFLOP/s (microbenchmark)", first["tflops"], "TFLOP/s", args.steps,
                     args.warmup, first["ms"], "per tier", "synthetic U[-1,1]",
                     {"workload": "tb_gemm NN", "m": M, "n": N, "k": K})
    line["tiers"] = res
    h.close()
    return line


def run_copy(args, torch, tensorblas):
    x = torch.rand(64 * 1024 * 1024, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / 3 * 1e3
    gbs = 2 * x.numel() * 4 / (ms * 1e-3) / 1e9
    return line_base("device copy GB/s", gbs, "GB/s", 3, 0, ms, "f32", "synthetic",
                     {"workload": "256 MiB device-to-device copy"})


# ------------------------------------------------------- the default bench line
# The other configurations of the benchmark, each measured in the same process after
# the headline TRSM and carried in the line under "configs": (key, workload, options).
DEFAULT_LEGS = [
    ("config1_syrk2048", "syrk", dict(n=2048, k=2048, tier="plain", steps=50, warmup=5)),
    ("syrk16384_split_vs_fp32", "syrk", dict(n=16384, k=16384, tier="split", steps=5, warmup=1)),
    ("config3_ozaki_syr2k_phi0.5", "ozaki", dict(op="syr2k", phi=0.5, steps=2, warmup=1)),
    ("config3_ozaki_syr2k_phi2", "ozaki", dict(op="syr2k", phi=2.0, steps=2, warmup=1)),
    ("config3_ozaki_symm_phi0.5", "ozaki", dict(op="symm", phi=0.5, steps=2, warmup=1)),
    ("config4_batched2000", "batched", dict(count=2000, tier="split", steps=10, warmup=2)),
]
LEG_DROP = ("n_gpus", "scaling", "vs_baseline", "higher_is_better")


def run_leg(fn, args, torch, tensorblas, **over):
    """One configuration of the default line: the workload's own line, or the error it
    raised (so that one failing leg does not hide the others)."""
    import copy
    a = copy.copy(args)
    for key in ("n", "m", "k", "count"):
        setattr(a, key, 0)
    a.tier, a.op, a.slices = "", "", ""
    for key, v in over.items():
        setattr(a, key, v)
    t0 = time.perf_counter()
    try:
        line = fn(a, torch, tensorblas)
        out = {k_: v for k_, v in line.items() if k_ not in LEG_DROP}
    except Exception as e:       # a library error status or a host-side failure
        out = {"error": f"{type(e).__name__}: {e}"}
    out["leg_seconds"] = round(time.perf_counter() - t0, 1)
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="trsm",
                    choices=["trsm", "syrk", "ozaki", "gemm", "copy", "batched", "serve", "potrf",
                             "trtri"])
    ap.add_argument("--slots", type=int, default=2)
    ap.add_argument("--budget-gib", type=float, default=32.0)
    ap.add_argument("--load", type=float, default=0.72)
    ap.add_argument("--large-ms", type=float, default=5.0)
    ap.add_argument("--max-large", type=int, default=1)
    ap.add_argument("--small-ms", type=float, default=0.0,
                    help="serve: estimated time alone below which a request is small")
    ap.add_argument("--small-slots", type=int, default=0,
                    help="serve: slots kept for small requests")
    ap.add_argument("--order", default="fifo", choices=["fifo", "shortest", "shortest-small"])
    ap.add_argument("--reserve-cus", type=int, default=0)
    ap.add_argument("--partition", type=int, default=0, choices=[0, 1, 2],
                    help="serve: --reserve-cus CUs left free of large requests, or with "
                         "--small-slots of every request outside the small lane (0: GEMM CU "
                         "caps, 1: CU-masked streams), or of the other requests (2: CU-masked "
                         "streams)")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="serve: HIP hardware queues of the process (default: one per slot, <= 32)")
    ap.add_argument("--count", type=int, default=0)
    ap.add_argument("--dump", default="", help="serve: write per-request records (JSON) here")
    ap.add_argument("--tier", default="")
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--m", type=int, default=0)
    ap.add_argument("--k", type=int, default=0)
    ap.add_argument("--nb", type=int, default=256)
    ap.add_argument("--tiers", default="plain,split,fp32")
    ap.add_argument("--slices", default="",
                    help="ozaki: slice counts to time, '+' = with the group g = S "
                         "(default: 8,8+,9 at phi 0.5; 8+,9+,10 at phi 2)")
    ap.add_argument("--phi", type=float, default=0.5)
    ap.add_argument("--op", default="", choices=["", "syr2k", "symm"])
    ap.add_argument("--split-planes", type=int, default=2, choices=[2, 3])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true",
                    help="trsm: skip the request-stream summary the default line carries")
    ap.add_argument("--no-configs", action="store_true",
                    help="trsm: skip the other configurations the default line carries")
    ap.add_argument("--cpu-budget", type=float, default=5.0,
                    help="batched: seconds of host BLAS for its cpu_baseline sample")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1")
    with_serve = args.workload == "trsm" and not args.no_serve
    if args.workload == "serve" or with_serve:
        # One hardware queue per slot stream, set before HIP starts.  With fewer queues
        # than slots HIP maps several slot streams onto one queue, and a small request's
        # kernels then wait in FIFO order behind a large request's on another slot.
        upper = args.slots - args.small_slots if args.small_ms > 0 and args.small_slots > 0 \
            else args.max_large
        extra = {0: 0, 1: upper, 2: args.slots}[args.partition]
        args.hw_queues = min(32, max(1, args.hw_queues or args.slots + extra))
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)

    import torch
    import tensorblas
    run = {"trsm": run_trsm, "syrk": run_syrk, "ozaki": run_ozaki, "gemm": run_gemm,
           "copy": run_copy, "batched": run_batched, "serve": run_serve,
           "potrf": run_potrf, "trtri": run_trtri}[args.workload]
    line = run(args, torch, tensorblas)
    if args.workload == "trsm" and not args.no_configs:
        fns = {"syrk": run_syrk, "ozaki": run_ozaki, "batched": run_batched}
        line["configs"] = {}
        for key, wl, over in DEFAULT_LEGS:
            log(f"bench: configuration {key}")
            line["configs"][key] = run_leg(fns[wl], args, torch, tensorblas, **over)
    if with_serve:
        # the metric's second half: the request stream (--workload serve, default
        # policy) in brief, measured after the TRSM in the same process
        torch.cuda.empty_cache()
        sv = run_serve(args, torch, tensorblas)
        line["served"] = {
            "workload": sv["config"]["workload"],
            "latency_ms": sv["latency_ms"], "isolated_latency_ms": sv["isolated_latency_ms"],
            "p99_latency_vs_isolated_p99": sv["p99_latency_vs_isolated_p99"],
            "latency_over_isolated": sv["latency_over_isolated"],
            "throughput_tflops": round(sv["value"], 3),
            "throughput_vs_offline": sv["throughput_vs_offline"],
            "results_equal_alone": sv["results_equal_alone"],
            "policy": {k: sv["config"][k] for k in ("slots", "load", "large_ms", "max_large", "order",
                                                   "reserve_cus", "partition", "small_ms",
                                                   "small_slots")},
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
