"""Benchmark of the SQUAREM EM hot path on one MI355X.

Workload (default): plain SQUAREM EM at human scale -- 200,000 transcripts,
2,000,000 drawn classes (singletons folded into unique counts, leaving
~1.3M multi-member classes, 4.56M members), 60M fragments, seed 0, uniform
start, stop at 1e-8 relative change or 1000 iterations.  One step is one
complete solve (src/EM.c:315-387) with the class table already resident on
the device; value = SQUAREM iterations per second over the timed steps.

Also config 3 (--bootstrap-reps, default 100): the bootstrap replicates in
one batched run, against the same replicates solved one by one on the GPU
and against the oracle on hera's own thread schedule; and config 4
(--grad-topk, default 1000): the adjoint gradient and Hessian-vector
products at the EM solution and the 50-step CG standard errors of the top
transcripts, with the oracle's CPU gradient timed beside it.

Modes (--modes, first one is the headline):
  fp64-exact  order-faithful FP64 E-step and the reference's sequential
              norms: every iterate bitwise the reference's;
  fp64-tree   the same E-step with fixed-shape parallel norms;
  fp32-tree   FP32 bulk E-step, FP64 M-step/SQUAREM, parallel norms.
Bootstrap modes (--boot-modes): fp64-exact, fp64-tree, fp32-exact (FP32
bulk E-step with the reference's FP64 norms), hera-sampler.
Each is checked against the FP64 CPU oracle on the same schedule, and the
oracle's own throughput on the host is the cpu_baseline.  The roofline is
the dominant memory-streaming kernel of the headline mode, timed with HIP
events on the solve's stream.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("EM iters/s + bootstrap reps/s @200k tx/2M classes vs unbatched "
          "GPU & CPU")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-tx", type=int, default=200000)
    ap.add_argument("--n-cls", type=int, default=2000000)
    ap.add_argument("--n-frag", type=int, default=60_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--stop-rel", type=float, default=1e-8,
                    help="relative-change stop threshold (hera's is 0.01)")
    ap.add_argument("--min-iter", type=int, default=50,
                    help="SQUAREM iterations always run (hera: 50)")
    ap.add_argument("--max-iter", type=int, default=1000,
                    help="iteration cap (hera: 1000); a smaller pair bounds "
                         "profiler runs")
    ap.add_argument("--modes", default="fp64-exact,fp64-tree,fp32-tree")
    ap.add_argument("--cpu-solves", type=int, default=1,
                    help="full oracle solves timed for cpu_baseline (0: skip)")
    ap.add_argument("--grad-topk", type=int, default=1000,
                    help="standard errors of this many transcripts (0: skip "
                         "the differentiation section)")
    ap.add_argument("--cg-steps", type=int, default=50)
    ap.add_argument("--bootstrap-reps", type=int, default=100,
                    help="bootstrap replicates of the batched run (0: skip)")
    ap.add_argument("--boot-modes",
                    default="fp64-exact,fp64-tree,fp32-exact,hera-sampler",
                    help="bootstrap modes to run")
    ap.add_argument("--seq-sample-fp32", type=int, default=10,
                    help="replicates solved one by one in the FP32 mode "
                         "(each runs to the iteration cap), scaled to all")
    ap.add_argument("--seq-sample", type=int, default=-1,
                    help="replicates solved one by one for the unbatched "
                         "comparison (-1: all of them; 0: skip)")
    ap.add_argument("--cpu-boot-threads", type=int, default=16,
                    help="threads (= replicates) of the CPU bootstrap sample")
    ap.add_argument("--no-tune", action="store_true",
                    help="use the default tile configuration instead of "
                         "the autotuner's")
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner cache file (default: none; tune now)")
    ap.add_argument("--skip-plain", action="store_true",
                    help="skip the plain solves (profiling runs of the other "
                         "sections); the JSON line then has no headline")
    ap.add_argument("--ragged-samples", type=int, default=64,
                    help="samples of the ragged multi-sample batch (0: skip)")
    ap.add_argument("--ragged-reps", type=int, default=20)
    ap.add_argument("--fallback-solves", type=int, default=3,
                    help="solves with the monotonicity fallback (0: skip)")
    ap.add_argument("--weighted-solves", type=int, default=3,
                    help="solves of the workload with member weights "
                         "(hera-EM's per-alignment probabilities); 0 skips")
    ap.add_argument("--ragged-seq-sample", type=int, default=8,
                    help="samples bootstrapped one handle at a time for the "
                         "comparison (scaled to all samples)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU "
              "path (bootstrap replicates and samples would shard across "
              "GPUs with no communication).  Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import hera_amd as H

    all_modes = {
        "fp64-exact": dict(precision=H.PREC_FP64, norm=H.NORM_EXACT),
        "fp64-exact-win": dict(precision=H.PREC_FP64,
                               norm=H.NORM_EXACT_WIN),
        "fp64-tree": dict(precision=H.PREC_FP64, norm=H.NORM_TREE),
        "fp32-tree": dict(precision=H.PREC_FP32, norm=H.NORM_TREE),
    }
    modes = {m: all_modes[m] for m in args.modes.split(",")}
    head_name = next(iter(modes))

    t0 = time.time()
    cls, u = H.generate(args.n_tx, args.n_cls, args.n_frag, seed=args.seed)
    log("generated T=%d C=%d Z=%d in %.2fs" % (cls.n_tx, cls.n_cls, cls.nnz,
                                               time.time() - t0))
    # the tile configuration: the autotuner's choice for this table and
    # replicate count (timed before, and outside, every measured region)
    cfg, tuning = None, None
    if not args.no_tune:
        t0 = time.time()
        cfg, rep = H.autotune(cls, u, n_rep=max(args.bootstrap_reps, 0),
                              iters=20, repeats=2, cache=args.tune_cache)
        tuning = dict(config=cfg, seconds=time.time() - t0, **rep)
        log("autotune %.1fs (%d candidates%s): %s; plain %.1f -> %.1f us/it, "
            "batched %.2f -> %.2f us per replicate-iteration" % (
                tuning["seconds"], rep["n_candidates"],
                ", cached" if rep["from_cache"] else "", cfg,
                rep["us_plain_default"], rep["us_plain_best"],
                rep["us_batch_default"], rep["us_batch_best"]))
    t0 = time.time()
    p = H.Problem(cls, config=cfg)
    t_upload = time.time() - t0
    log("uploaded class table in %.3fs" % t_upload)

    sk = dict(stop_rel_change=args.stop_rel, min_iter=args.min_iter,
              max_iter=args.max_iter)
    if args.skip_plain:
        out = {"metric": METRIC, "skipped": "plain solves (--skip-plain)",
               "tuning": tuning}
        # the copy of known size that calibrates profiler byte counters
        out["copy_bandwidth_GBps"] = H.copy_bandwidth(1 << 30, 10)
        if args.bootstrap_reps > 0:
            boot = bench_bootstrap(args, H, np, p, cls, u, sk)
            boot.pop("_est_exact", None)
            for v in boot.values():
                if isinstance(v, dict):
                    v.pop("_iters", None)
            out["bootstrap"] = boot
        if args.ragged_samples > 0:
            out["ragged"] = bench_ragged(args, H, np)
        p.close()
        print(json.dumps(out), flush=True)
        return
    results = {}
    head_result = None
    for name, mk in modes.items():
        steps = args.steps if name == head_name else max(args.steps // 2, 1)
        for _ in range(args.warmup):
            p.solve(u, **mk, **sk)
        it = 0
        loop_ms = 0.0
        r = None
        t = time.perf_counter()
        for _ in range(steps):
            r = p.solve(u, **mk, **sk)
            it += r["stats"]["iters"]
            if name == head_name:
                head_result = r
            loop_ms += r["stats"]["ms_solve"]
        dt = time.perf_counter() - t
        # value: whole calls (unique counts up, tpm/est/lambda down, setup
        # and finalisation included); loop: HIP events around the
        # iteration graphs alone, table and counts resident
        results[name] = dict(steps=steps, iters=it, seconds=dt,
                             iters_per_s=it / dt, ms_per_solve=dt / steps * 1e3,
                             iters_per_solve=r["stats"]["iters"],
                             loop_iters_per_s=it / (loop_ms * 1e-3),
                             exact_sum_reruns=r["stats"]["sum_reruns"])
        log("%-10s %d solves %.4fs  %.1f it/s  k=%d" % (
            name, steps, dt, it / dt, r["stats"]["iters"]))

    head = results[head_name]
    prof = p.profile(u, **modes[head_name], **sk)
    kinds = [k for k in prof if k != "iters"]
    # the roofline names the dominant memory-streaming kernel; the exact
    # sums' in-order chunk walk is latency-bound by construction
    stream_kinds = [k for k in kinds if k not in ("norm_pass1", "norm_pass2")
                    or modes[head_name]["norm"] not in (H.NORM_EXACT,
                                                        H.NORM_EXACT_WIN)]
    dom = max(stream_kinds, key=lambda k: prof[k]["ms_total"])
    d = prof[dom]
    # algorithmic bytes: 8Z + 16C + 24T per EM map (u32 indices, FP32
    # values: the compulsory traffic of a class pass and a transcript pass
    # with the M-step), times the maps one launch applies
    b_map = map_bytes(cls)
    maps = MAPS_PER_LAUNCH.get(dom)
    alg = b_map * maps if maps else d["bytes_per_launch"]
    achieved = alg / (d["ms_avg"] * 1e-3) / 1e9
    tot_ms = sum(prof[k]["ms_total"] for k in kinds)
    for k in kinds:
        if prof[k]["launches"]:
            log("kernel %-16s launches=%5d avg=%.4f ms (%4.1f%%) %.1f GB/s" % (
                k, prof[k]["launches"], prof[k]["ms_avg"],
                100 * prof[k]["ms_total"] / tot_ms,
                prof[k]["bytes_per_launch"] / (prof[k]["ms_avg"] * 1e-3) / 1e9))
    it_ms = tot_ms / max(prof["iters"], 1)

    out = {
        "metric": METRIC,
        "value": head["iters_per_s"],
        "unit": "EM iterations/s (SQUAREM cycles of 3 EM maps)",
        "n_gpus": 1,
        "steps": head["steps"],
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_solve"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if modes[head_name]["precision"] == H.PREC_FP64
                 else "f32 E-step / f64 M-step+SQUAREM",
        "data": "synthetic",
        "config": {
            "workload": "plain SQUAREM EM, one solve per step, mode %s"
                        % head_name,
            "n_tx": cls.n_tx, "n_cls_drawn": args.n_cls,
            "n_cls_multi": cls.n_cls, "nnz": cls.nnz,
            "n_frag": args.n_frag, "seed": args.seed,
            "iters_per_solve": head["iters_per_solve"],
            "stop_rule": "min 50, max 1000 iterations, relative change %g "
                         "on lambda > 1e-3 (hera's test, src/EM.c:292-297)"
                         % args.stop_rel,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profile(dom, args),
            "algorithmic_bytes_per_launch": alg,
            "bytes_rule": "8Z + 16C + 24T per EM map (u32 indices, FP32 "
                          "values) x %s maps per launch" % maps,
            "kernel_own_bytes_per_launch": d["bytes_per_launch"],
            "ms_per_launch": d["ms_avg"],
            "iteration": {
                "ms": it_ms,
                "algorithmic_bytes": 3 * b_map + 80 * cls.n_tx,
                "achieved_GBps": (3 * b_map + 80 * cls.n_tx)
                                 / (it_ms * 1e-3) / 1e9,
                "note": "one SQUAREM iteration, all its kernels, by HIP "
                        "events: 3 maps + 80T of FP64 vector traffic"},
        },
        "modes": results,
        "kernels_headline": {k: dict(prof[k]) for k in kinds
                             if prof[k]["launches"]},
        "upload_s": t_upload,
        "tuning": tuning,
    }

    out["copy_bandwidth_GBps"] = H.copy_bandwidth(1 << 30, 10)
    log("device copy (1 GiB, read+write): %.0f GB/s"
        % out["copy_bandwidth_GBps"])

    out["writer"] = bench_writer(H, np, cls, u, head_result)

    if args.grad_topk > 0:
        out["differentiation"] = bench_grad(args, H, np, p, cls, u, sk)

    if args.ragged_samples > 0:
        out["ragged"] = bench_ragged(args, H, np)

    if args.weighted_solves > 0:
        out["weighted"] = bench_weighted(args, H, np, cls, u, sk)

    if args.fallback_solves > 0:
        out["fallback"] = bench_fallback(args, H, p, u, sk, head_result)

    boot = None
    if args.bootstrap_reps > 0:
        boot = bench_bootstrap(args, H, np, p, cls, u, sk)
        out["bootstrap"] = boot

    if args.cpu_solves > 0 or args.cpu_boot_threads > 0:
        from oracle import oracle as O
    if args.cpu_solves > 0:
        n = args.cpu_solves
        t0 = O.now()
        k_cpu = 0
        ora = None
        for _ in range(n):
            ora = O.estimate_count(cls, u, stop_rel=args.stop_rel,
                                   min_round=args.min_iter,
                                   max_round=args.max_iter)
            k_cpu += ora["k"]
        dt_cpu = O.now() - t0
        out["cpu_baseline"] = dict({
            "value": k_cpu / dt_cpu,
            "unit": "EM iterations/s",
            "cores": 1,
            "kind": "port",
            "sample": "%d full plain solve(s) of this workload (%d iterations, "
                      "%.1fs) by the FP64 C restatement of src/EM.c, "
                      "gcc -O2, 1 thread" % (n, k_cpu, dt_cpu),
        }, **cpu_info())
        m = ora["tpm"] / 1e6 > 1e-7
        ll_o = O.loglik(cls, u, ora["lam"])
        par = {"k_oracle": ora["k"], "abund_threshold": "tpm/1e6 > 1e-7"}
        for name, mk in modes.items():
            g = p.solve(u, **mk, **dict(sk, min_iter=ora["k"],
                                        max_iter=ora["k"]))
            e = np.abs(g["tpm"][m] - ora["tpm"][m]) / ora["tpm"][m]
            par[name] = {
                "k_gpu_stop_rule": results[name]["iters_per_solve"],
                "lambda_bitwise_equal": bool(np.array_equal(g["lam"],
                                                            ora["lam"])),
                "abund_rel_err_max": float(e.max()),
                "abund_rel_err_median": float(np.median(e)),
                "abund_frac_over_1e-6": float(np.mean(e > 1e-6)),
                "loglik_rel_err": abs(g["stats"]["loglik"] - ll_o) / abs(ll_o),
            }
        out["parity"] = par
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    if boot is not None:
        if args.cpu_boot_threads > 0:
            bootstrap_cpu(args, np, O, cls, u, boot)
        boot.pop("_est_exact", None)
        for v in boot.values():
            if isinstance(v, dict):
                v.pop("_iters", None)
    p.close()
    print(json.dumps(out), flush=True)


MAPS_PER_LAUNCH = {"tile_map12": 2, "tile_ex3_map12": 3, "class_pass": None,
                   "transcript_pass": None}


def map_bytes(cls):
    """Algorithmic bytes of one EM map: 8Z + 16C + 24T (row_ptr 4C, col 4Z,
    n 4C, rho write 4C, lambda 4T; col_ptr 4T, row 4Z, rho read 4C, lambda
    4T; M-step u 4T, w 4T, lambda' 4T)."""
    return 8.0 * cls.nnz + 16.0 * cls.n_cls + 24.0 * cls.n_tx


def batch_map_bytes(cls, reps):
    """The same for one map of `reps` replicates sharing the structure:
    8Z + 4C + 8T once, 12C + 16T per replicate."""
    return (8.0 * cls.nnz + 4.0 * cls.n_cls + 8.0 * cls.n_tx
            + reps * (12.0 * cls.n_cls + 16.0 * cls.n_tx))


def cpu_info():
    """The host's processor: model, clock, logical CPUs and the CPUs this
    process may run on."""
    model, mhz = None, None
    try:
        for line in open("/proc/cpuinfo"):
            if model is None and line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
            if mhz is None and line.startswith("cpu MHz"):
                mhz = float(line.split(":", 1)[1])
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count()
    return {"cpu_model": model, "cpu_mhz": mhz, "nproc": os.cpu_count(),
            "cpus_available": aff}


def traffic_from_profile(kind, args):
    """HBM bytes per launch of the headline kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated) of the
    default workload, or None for any other workload."""
    if (args.n_tx, args.n_cls, args.n_frag, args.seed) != (
            200000, 2000000, 60_000_000, 0):
        return None
    kernel = {"tile_map12": "k_ptile_map12",
              "tile_ex3_map12": "k_ptile_ex3_map12"}.get(kind)
    path = os.path.join(ROOT, "profiles",
                        "em_kernels_200k_tx_2M_cls_counters.md")
    if kernel is None or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch %s:" % kernel):
            return float(line.split(":")[1])
    return None


def batch_traffic_from_profile(args, R, mode):
    """Mean HBM bytes per launch of the batched tile kernel over the
    bootstrap run, from the committed copy-calibrated FETCH_SIZE /
    WRITE_SIZE summary of the default workload (exact mode, 100
    replicates), or None for any other run."""
    if (args.n_tx, args.n_cls, args.n_frag, args.seed, R, mode) != (
            200000, 2000000, 60_000_000, 0, 100, "fp64-exact"):
        return None
    path = os.path.join(ROOT, "profiles",
                        "bootstrap_tile_ex3_200k_tx_2M_cls_R100_counters.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch k_tile_ex3:"):
            return float(line.split(":")[1])
    return None


def bench_grad(args, H, np, p, cls, u, sk):
    """Config 4: the adjoint at the EM solution (theta = log(lambda len),
    log_efflen = log len), Hessian-vector products, and the CG standard
    errors of the top transcripts; the oracle's CPU gradient is timed beside
    it.  Accuracy against central differences is covered by the tests."""
    from oracle import oracle as O
    g = p.solve(u, norm=H.NORM_TREE, **sk)
    lam = g["lam"]
    with np.errstate(divide="ignore"):
        theta = np.log(lam * cls.tx_len)
    theta -= theta[np.isfinite(theta)].max()
    ell = np.log(cls.tx_len.astype(float))
    reps = 10
    res = {}
    b_map = map_bytes(cls)
    for name, prec in (("fp64", H.PREC_FP64), ("fp32_forward", H.PREC_FP32)):
        p.grad(u, theta, ell, precision=prec)
        ms, ma = [], []
        for _ in range(reps):
            r = p.grad(u, theta, ell, precision=prec)
            ms.append(r["ms"])
            ma.append(r["ms_adjoint"])
        res["grad_" + name + "_ms"] = float(np.median(ms))
        res["grad_" + name + "_adjoint_ms"] = float(np.median(ma))
    # the adjoint's class and transcript passes read what one EM map reads
    # (a class pass and a transcript pass over the same table)
    ach = b_map / (res["grad_fp64_adjoint_ms"] * 1e-3) / 1e9
    res["roofline"] = {
        "bound": "hbm", "kernel": "k_ptile_grad (adjoint passes)",
        "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": ach / HBM_PEAK_GBS,
        "algorithmic_bytes": b_map,
        "bytes_rule": "8Z + 16C + 24T (one class pass and one transcript "
                      "pass, u32 indices, FP32 values)",
        "ms": res["grad_fp64_adjoint_ms"]}
    gg = p.grad(u, theta, ell)
    E = -gg["g_log_efflen"]
    res["grad_rel_stationarity"] = float(
        np.max(np.abs(gg["g_theta"]) / (E + 1.0)))
    v = np.random.default_rng(0).normal(size=cls.n_tx)
    p.hvp(u, theta, ell, v)
    t0 = time.perf_counter()
    for _ in range(reps):
        p.hvp(u, theta, ell, v)
    res["hvp_wall_ms"] = (time.perf_counter() - t0) / reps * 1e3
    p.stderr(u, theta, ell, top_k=args.grad_topk, cg_steps=2)
    se = p.stderr(u, theta, ell, top_k=args.grad_topk,
                  cg_steps=args.cg_steps)
    st = se["stats"]
    res.update({
        "top_k": args.grad_topk, "cg_steps": args.cg_steps,
        "se_ms_setup": st["ms_setup"], "se_ms_cg": st["ms_cg"],
        "se_ms_total": st["ms_total"], "se_resid_max": st["resid_max"],
        "se_resid_median": float(np.median(se["resid"])),
        "se_chunk": st["chunk"],
        "se_of_top1_tpm": float(se["se"][0] * 1e6),
    })
    t0 = O.now()
    O.grad(cls, u, np.where(np.isfinite(theta), theta, -745.0), ell)
    res["cpu_grad_ms"] = (O.now() - t0) * 1e3
    res["cpu_grad_note"] = "oracle FP64 gradient, 1 thread"
    log("adjoint passes %.4f ms: %.0f GB/s by algorithmic bytes (%.1f%% "
        "of peak)" % (res["grad_fp64_adjoint_ms"], ach,
                      100 * ach / HBM_PEAK_GBS))
    log("grad fp64 %.3f ms, fp32-forward %.3f ms, HVP %.2f ms (wall), "
        "CPU grad %.1f ms; SE top %d, %d CG steps: %.1f ms (resid max %.1e)"
        % (res["grad_fp64_ms"], res["grad_fp32_forward_ms"],
           res["hvp_wall_ms"], res["cpu_grad_ms"], args.grad_topk,
           args.cg_steps, st["ms_cg"], st["resid_max"]))
    return res


def bench_writer(H, np, cls, u, res):
    """hera's abundance.tsv and abundance.gene.tsv for this solve's result
    (host formatting and file writes; names and genes synthetic)."""
    import tempfile
    T = cls.n_tx
    names = ["SYNT%011d.1|tx%d" % (i, i) for i in range(T)]
    gene = (np.arange(T) // 4).astype(np.uint32)
    gnames = ["SYNG%011d.1|g%d" % (g, g) for g in range(int(gene[-1]) + 1)]
    eff = cls.tx_len.astype(np.float64)
    with tempfile.TemporaryDirectory() as d:
        t = time.perf_counter()
        H.write_result(d, "bench", names, u, cls.tx_len, eff, res["est"],
                       res["tpm"], gene=gene, gene_names=gnames)
        dt = time.perf_counter() - t
        size = sum(os.path.getsize(os.path.join(d, f))
                   for f in os.listdir(d))
    log("writer: %d rows, %.1f ms, %.1f MB" % (T, dt * 1e3, size / 1e6))
    return {"rows": T, "ms": dt * 1e3, "bytes": size,
            "note": "hem_write_result: abundance.tsv + abundance.gene.tsv "
                    "(4 transcripts per gene), including packing the name "
                    "blocks in Python"}


def bench_fallback(args, H, p, u, sk, plain):
    """The monotonicity fallback (hem_opts.fallback) on the headline table:
    FP64 exact solves whose iterations end at u2 when the extrapolated map
    lowers the log-likelihood, whole calls timed; the final log-likelihood
    beside the plain solve's."""
    mk = dict(precision=H.PREC_FP64, norm=H.NORM_EXACT, fallback=1)
    p.solve(u, **mk, **sk)
    it = 0
    t = time.perf_counter()
    for _ in range(args.fallback_solves):
        r = p.solve(u, **mk, **sk)
        it += r["stats"]["iters"]
    dt = time.perf_counter() - t
    st = r["stats"]
    res = dict(solves=args.fallback_solves, iters_per_solve=st["iters"],
               iters_per_s=it / dt,
               ms_per_solve=dt / args.fallback_solves * 1e3,
               fallbacks_per_solve=st["fallbacks"], loglik=st["loglik"],
               loglik_plain=plain["stats"]["loglik"] if plain else None,
               iters_plain=plain["stats"]["iters"] if plain else None)
    log("fallback  %d solves %.4fs  %.1f it/s  k=%d (%d fell back); loglik "
        "%.12e vs plain %.12e" % (args.fallback_solves, dt, it / dt,
                                  st["iters"], st["fallbacks"], st["loglik"],
                                  res["loglik_plain"] or 0.0))
    return res


def bench_weighted(args, H, np, cls, u, sk):
    """The same table with a weight per member (uniform in [0.05, 1),
    seeded): FP64 exact solves on the order-faithful class and transcript
    passes, whole calls timed, plus the per-kernel profile of one solve."""
    import dataclasses
    rng = np.random.default_rng(args.seed + 1)
    wc = dataclasses.replace(cls, member_weight=rng.uniform(0.05, 1.0,
                                                            cls.nnz))
    pw = H.Problem(wc)
    mk = dict(precision=H.PREC_FP64, norm=H.NORM_EXACT)
    pw.solve(u, **mk, **sk)
    it = 0
    t = time.perf_counter()
    for _ in range(args.weighted_solves):
        r = pw.solve(u, **mk, **sk)
        it += r["stats"]["iters"]
    dt = time.perf_counter() - t
    prof = pw.profile(u, **mk, **sk)
    pw.close()
    kern = {}
    for k, d in prof.items():
        if k == "iters" or not d["launches"]:
            continue
        kern[k] = dict(launches=d["launches"], ms_avg=d["ms_avg"],
                       bytes_per_launch=d["bytes_per_launch"],
                       GBps=d["bytes_per_launch"] / (d["ms_avg"] * 1e-3)
                       / 1e9)
    res = dict(solves=args.weighted_solves, iters_per_solve=r["stats"]["iters"],
               iters_per_s=it / dt, ms_per_solve=dt / args.weighted_solves
               * 1e3, kernels=kern,
               weights="uniform [0.05, 1) per member, seed %d" % (args.seed + 1))
    log("weighted  %d solves %.4fs  %.1f it/s  k=%d" % (
        args.weighted_solves, dt, it / dt, r["stats"]["iters"]))
    return res


def bench_ragged(args, H, np):
    """Config 5: samples with transcript counts log-uniform in
    5,000-200,000, 10x as many drawn classes, fragments uniform in 5M-80M,
    each with R bootstrap replicates, all in one multi-sample handle (every
    (sample, replicate) EM instance in the same launches, per-instance stop
    tests); against the same samples bootstrapped one handle at a time, and
    every sample's gradient in one call at its first replicate's
    estimate."""
    S, R = args.ragged_samples, args.ragged_reps
    rng = np.random.default_rng(12345)
    ntx = np.exp(rng.uniform(np.log(5000), np.log(200000), S)).astype(int)
    nfr = rng.integers(5_000_000, 80_000_001, S)
    t0 = time.time()
    gens = [H.generate(int(t), int(10 * t), int(f), seed=1000 + q)
            for q, (t, f) in enumerate(zip(ntx, nfr))]
    cls = [g[0] for g in gens]
    us = [g[1] for g in gens]
    log("ragged: %d samples, %d transcripts, %d classes, %d members "
        "generated in %.1fs" % (S, sum(c.n_tx for c in cls),
                                sum(c.n_cls for c in cls),
                                sum(c.nnz for c in cls), time.time() - t0))
    seeds = (np.arange(S * R, dtype=np.uint64) + 1).reshape(S, R)
    sk = dict(stop_rel_change=args.stop_rel, min_iter=args.min_iter,
              max_iter=args.max_iter)
    t0 = time.time()
    mp = H.Problem.from_samples(cls)
    res = {"n_samples": S, "reps_per_sample": R,
           "n_tx_total": int(mp.n_tx), "n_cls_total": int(mp.cls0[-1]),
           "nnz_total": int(sum(c.nnz for c in cls)),
           "upload_s": time.time() - t0, "seeds": "1..%d" % (S * R)}
    u_all = np.concatenate(us)
    out_multi = {}
    for name, nm in (("fp64-tree", H.NORM_TREE),
                     ("fp64-exact", H.NORM_EXACT)):
        mp.bootstrap(u_all, seeds[:, :2], norm=nm, **sk)   # graphs
        t = time.perf_counter()
        b = mp.bootstrap(u_all, seeds, norm=nm, **sk)
        dt = time.perf_counter() - t
        st = b["stats"]
        res[name] = dict(seconds=dt, instances_per_s=S * R / dt,
                         ms_resample=st["ms_resample"],
                         ms_solve=st["ms_solve"],
                         iters_min=int(b["iters"].min()),
                         iters_max=int(b["iters"].max()),
                         iters_total=int(b["iters"].sum()))
        out_multi[name] = b
        log("ragged %-10s %d samples x %d reps: %.3fs (%.0f instances/s) "
            "iters %d..%d" % (name, S, R, dt, S * R / dt,
                              b["iters"].min(), b["iters"].max()))
    # a gradient pass for every sample in one call on the same handle, at
    # each sample's first exact replicate (theta = log est, softmax within
    # the sample), against the samples' own handles one at a time
    est0 = out_multi["fp64-exact"]["est"][0]
    with np.errstate(divide="ignore"):
        th_all = np.log(est0)
    th_all = np.where(np.isfinite(th_all), th_all, -745.0)
    ell_all = np.log(np.concatenate([c.tx_len for c in cls]).astype(float))
    mp.grad_samples(u_all, th_all, ell_all)
    gm = [mp.grad_samples(u_all, th_all, ell_all) for _ in range(5)]
    mp.close()
    g_ms = float(np.median([g["ms"] for g in gm]))
    g_ma = float(np.median([g["ms_adjoint"] for g in gm]))
    tg, worst = 0.0, 0.0
    for q in range(S):
        a, e = int(mp.tx0[q]), int(mp.tx0[q + 1])
        p = H.Problem(cls[q])
        p.grad(us[q], th_all[a:e], ell_all[a:e])
        own = p.grad(us[q], th_all[a:e], ell_all[a:e])
        p.close()
        tg += own["ms"]
        worst = max(worst, float(np.max(np.abs(gm[0]["g_theta"][a:e]
                                               - own["g_theta"]))
                                 / np.max(np.abs(own["g_theta"]))))
    res["gradient"] = {
        "one_call_ms": g_ms, "adjoint_ms": g_ma,
        "own_handles_ms_total": tg, "speedup": tg / g_ms,
        "max_rel_diff_vs_own_handles": worst,
        "note": "hem_grad_samples on the multi-sample handle (device ms, "
                "uploads included) against hem_grad on each sample's own "
                "handle, summed"}
    log("ragged gradient: one call %.2f ms (adjoint %.2f ms) vs %.2f ms on "
        "%d own handles; max rel diff %.1e" % (g_ms, g_ma, tg, S, worst))
    # the same samples one handle at a time (a sample of them, scaled)
    nseq = min(args.ragged_seq_sample, S)
    pick = np.linspace(0, S - 1, nseq).astype(int)
    for name, nm in (("fp64-tree", H.NORM_TREE),):
        tsum = 0.0
        same = True
        for q in pick:
            p = H.Problem(cls[q])
            p.bootstrap(us[q], seeds[q, :2], norm=nm, **sk)
            t = time.perf_counter()
            b = p.bootstrap(us[q], seeds[q], norm=nm, **sk)
            tsum += time.perf_counter() - t
            p.close()
            t0q = int(sum(c.n_tx for c in cls[:q]))
            same &= bool(np.array_equal(
                out_multi[name]["est"][:, t0q:t0q + cls[q].n_tx], b["est"]))
        frac = sum(cls[q].n_tx for q in pick) / sum(c.n_tx for c in cls)
        est_all = tsum / frac
        res[name]["one_handle_at_a_time_s"] = est_all
        res[name]["speedup_vs_one_at_a_time"] = est_all / res[name]["seconds"]
        res[name]["bitwise_equal_own_handles"] = same
        log("ragged %s one handle at a time: %.3fs for %d samples (%.1f%% of "
            "transcripts) -> %.2fs for all; x%.2f; bitwise %s" % (
                name, tsum, nseq, 100 * frac, est_all,
                res[name]["speedup_vs_one_at_a_time"], same))
    return res


def bench_bootstrap(args, H, np, p, cls, u, sk):
    """Config 3: R bootstrap replicates (multinomial over all fragments, one
    seed per replicate) in one batched run, against the same replicates
    solved one after another by the plain GPU path."""
    R = args.bootstrap_reps
    seeds = np.arange(R, dtype=np.uint64) + 1
    res = {"n_rep": R, "seeds": "1..%d" % R}
    # fp32-exact: the FP32 bulk E-step with the reference's FP64 norms
    # (the precision form of the batched workload); every replicate runs
    # to the iteration cap, since FP32 E-step noise stays above a 1e-8
    # relative change
    allm = (("fp64-exact", dict(norm=H.NORM_EXACT)),
            ("fp64-exact-win", dict(norm=H.NORM_EXACT_WIN)),
            ("fp64-tree", dict(norm=H.NORM_TREE)),
            ("fp32-exact", dict(norm=H.NORM_EXACT, precision=H.PREC_FP32)))
    bmodes = [(n, m) for n, m in allm if n in args.boot_modes.split(",")]
    for name, nm in bmodes:
        p.bootstrap(u, seeds[:2], **nm, **sk)   # capture the graphs
        t = time.perf_counter()
        b = p.bootstrap(u, seeds, **nm, **sk)
        dt = time.perf_counter() - t
        st = b["stats"]
        res[name] = dict(reps_per_s=R / dt, seconds=dt,
                         ms_resample=st["ms_resample"],
                         ms_solve=st["ms_solve"],
                         iters_min=st["iters_min"], iters_max=st["iters_max"],
                         iters_total=int(b["iters"].sum()),
                         chunk=st["chunk"], n_tiles=st["n_tiles"],
                         exact_sum_reruns=st["sum_reruns"])
        res[name]["_iters"] = b["iters"]
        if name == "fp64-exact":
            res["_est_exact"] = b["est"]
        log("bootstrap %-10s R=%d %.3fs (%.1f reps/s) resample %.1f ms "
            "iters %d..%d" % (name, R, dt, R / dt, st["ms_resample"],
                              st["iters_min"], st["iters_max"]))
    if "hera-sampler" in args.boot_modes.split(","):
        # hera's own sampler (unique reads redrawn with rand_r, class
        # counts kept), the replicates of a one-thread hera run (seed 0)
        p.bootstrap_hera(u, 2, 0, **sk)
        t = time.perf_counter()
        b = p.bootstrap_hera(u, R, 0, **sk)
        dt = time.perf_counter() - t
        st = b["stats"]
        res["fp64-exact-hera-sampler"] = dict(
            reps_per_s=R / dt, seconds=dt, ms_resample=st["ms_resample"],
            ms_solve=st["ms_solve"], iters_min=st["iters_min"],
            iters_max=st["iters_max"], iters_total=int(b["iters"].sum()),
            unique_reads=int(u.sum()), seed=0)
        log("bootstrap hera-sampler R=%d %.3fs (%.1f reps/s) resample "
            "%.1f ms iters %d..%d" % (R, dt, R / dt, st["ms_resample"],
                                      st["iters_min"], st["iters_max"]))
        if args.cpu_boot_threads > 0:
            from oracle import oracle as O
            t = time.perf_counter()
            cu, _ = O.hera_sample(u, 1, 0)
            dtc = time.perf_counter() - t
            ok = bool(np.array_equal(cu[0],
                                     p.bootstrap_hera_counts(u, 1, 0)[0]))
            res["fp64-exact-hera-sampler"]["cpu_sampler_ms_per_rep"] = \
                dtc * 1e3
            res["fp64-exact-hera-sampler"]["gpu_sampler_ms_per_rep"] = \
                st["ms_resample"] / R
            res["fp64-exact-hera-sampler"]["counts_bitwise_vs_oracle"] = ok
            log("hera sampler: CPU %.1f ms per replicate (1 thread), GPU "
                "%.2f ms; counts bitwise %s" % (dtc * 1e3,
                                                st["ms_resample"] / R, ok))
    # unbatched: the same resampled counts through the plain solve, one
    # replicate after another (all R unless --seq-sample asks for fewer)
    if args.seq_sample == 0:
        return res
    n_all = R if args.seq_sample < 0 else min(args.seq_sample, R)
    uo, co = p.bootstrap_counts(u, seeds)
    for name, nm in bmodes:
        n_seq = n_all
        if name.startswith("fp32"):
            n_seq = max(1, min(n_all, args.seq_sample_fp32))
        its = []
        t = time.perf_counter()
        for r in range(n_seq):
            p.set_counts(co[r])
            its.append(p.solve(uo[r], **nm, **sk)["stats"]["iters"])
        dt = (time.perf_counter() - t) * R / n_seq
        its = np.array(its)
        res[name]["sequential_plain_s"] = dt
        res[name]["sequential_plain_reps_per_s"] = R / dt
        res[name]["speedup_vs_sequential_gpu"] = dt / res[name]["seconds"]
        res[name]["sequential_replicates_timed"] = n_seq
        res[name]["sequential_iters"] = {
            "min": int(its.min()), "max": int(its.max()),
            "total": int(its.sum()),
            "same_as_batched": bool(np.array_equal(
                its, res[name]["_iters"][:n_seq]))}
        log("sequential %-10s %d replicates one by one: %.3fs (x%.2f); "
            "iters %d..%d total %d" % (name, n_seq, dt,
                                       res[name]["speedup_vs_sequential_gpu"],
                                       its.min(), its.max(), its.sum()))
    p.set_counts(cls.cls_count)
    res["sequential_sample"] = (
        "the first n of %d replicates (sequential_replicates_timed per "
        "mode) solved one by one (hem_problem_set_counts, then hem_solve), "
        "whole calls timed, scaled to all" % R)
    # the batched kernels launch by launch (HIP events) on the same counts
    for name, nm in bmodes:
        pr = p.profile_batch(uo, co, **nm, **sk)
        kern = {}
        for k in ("tile_map12", "norm_pass1", "norm_pass2", "tile_ex3",
                  "iter_end"):
            d = pr[k]
            if d["launches"]:
                kern[k] = dict(d, GBps=d["bytes_total"]
                               / (d["ms_total"] * 1e-3) / 1e9)
        ex = pr["tile_ex3"]
        ach = ex["bytes_total"] / (ex["ms_total"] * 1e-3) / 1e9
        res[name]["kernels"] = kern
        res[name]["roofline"] = {
            "bound": "hbm", "kernel": "k_tile_ex3",
            "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS,
            "traffic": batch_traffic_from_profile(args, R, name),
            "bytes_rule": "per launch 3 maps x (8Z + 4C + 8T + R_active "
                          "(12C + 16T)), u32 indices and FP32 values, "
                          "summed over launches",
            "ms_per_replicate_iteration": ex["ms_total"] / pr["rep_iters"],
            "launches": ex["launches"], "rep_iters": pr["rep_iters"]}
        log("batched %-10s k_tile_ex3 %d launches %.1f ms: %.0f GB/s by "
            "algorithmic bytes (%.1f%% of peak), %.1f us per "
            "replicate-iteration" % (
                name, ex["launches"], ex["ms_total"], ach,
                100 * ach / HBM_PEAK_GBS,
                1e3 * ex["ms_total"] / pr["rep_iters"]))
    return res


def bootstrap_cpu(args, np, O, cls, u, boot):
    """The oracle on hera's own bootstrap schedule (batches of n threads,
    joined per batch; each thread resamples and solves), on one batch of
    as many replicates as threads.  Threads: the CPUs this process may use,
    at most --cpu-boot-threads (16, the host share of one GPU on the
    benchmark machines)."""
    info = cpu_info()
    nthr = max(1, min(args.cpu_boot_threads, info["cpus_available"] or 1))
    seeds = np.arange(nthr, dtype=np.uint64) + 1
    t0 = O.now()
    est, ks = O.bootstrap_resampled(cls, u, seeds, nthr,
                                    stop_rel=args.stop_rel,
                                    min_round=args.min_iter,
                                    max_round=args.max_iter)
    dt = O.now() - t0
    boot["cpu_baseline"] = dict({
        "value": nthr / dt, "unit": "bootstrap replicates/s",
        "cores": nthr, "kind": "port",
        "sample": "one batch of %d replicates (seeds 1..%d, iterations "
                  "%d..%d) on %d threads: resample + FP64 restatement of "
                  "src/EM.c, %.1fs" % (nthr, nthr, int(ks.min()),
                                       int(ks.max()), nthr, dt)}, **info)
    gpu = boot.pop("_est_exact", None)
    same = [gpu is not None and bool(np.array_equal(gpu[r], est[r]))
            for r in range(nthr)]
    if gpu is not None:
        boot["parity_exact_vs_oracle"] = {"replicates": nthr,
                                          "est_bitwise_equal": all(same)}
    for name in ("fp64-exact", "fp64-tree", "fp32-exact"):
        if name in boot:
            boot[name]["speedup_vs_cpu"] = (boot[name]["reps_per_s"]
                                            / (nthr / dt))
    log("bootstrap cpu: %d reps on %d threads %.1fs; exact bitwise %s" % (
        nthr, nthr, dt, all(same)))


if __name__ == "__main__":
    main()
