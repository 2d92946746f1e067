"""Benchmark of the SQUAREM EM hot path on one MI355X.

Workload (default): plain SQUAREM EM at human scale -- 200,000 transcripts,
2,000,000 drawn classes (singletons folded into unique counts, leaving
~1.3M multi-member classes, 4.56M members), 60M fragments, seed 0, uniform
start, stop at 1e-8 relative change or 1000 iterations.  The timed mode is
the FP64 parity mode (order-faithful E-step); the FP32-bulk throughput mode
is timed beside it.  One step is one complete solve (src/EM.c:315-387) with
the class table already resident on the device; value = SQUAREM iterations
per second summed over the timed steps.

Also reported: the per-kernel roofline of the dominant kernel (HIP events on
the solve's stream), the parity of the timed solve against the FP64 CPU
oracle, and the oracle's own throughput on the host (cpu_baseline).
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
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-tx", type=int, default=200000)
    ap.add_argument("--n-cls", type=int, default=2000000)
    ap.add_argument("--n-frag", type=int, default=60_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp64")
    ap.add_argument("--stop-rel", type=float, default=1e-8,
                    help="relative-change stop threshold (hera's is 0.01)")
    ap.add_argument("--cpu-solves", type=int, default=2,
                    help="full oracle solves timed for cpu_baseline (0: skip)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU "
              "path (replicates and problems would shard with no "
              "communication; see DESIGN.md).  Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import hera_amd as H

    prec = H.PREC_FP32 if args.precision == "fp32" else H.PREC_FP64
    t0 = time.time()
    cls, u = H.generate(args.n_tx, args.n_cls, args.n_frag, seed=args.seed)
    log("generated T=%d C=%d Z=%d in %.2fs" % (cls.n_tx, cls.n_cls, cls.nnz,
                                               time.time() - t0))
    t0 = time.time()
    p = H.Problem(cls)
    t_upload = time.time() - t0
    log("uploaded class table in %.3fs" % t_upload)

    sk = dict(stop_rel_change=args.stop_rel)

    def timed(pr, steps, warmup):
        for _ in range(warmup):
            p.solve(u, precision=pr, **sk)
        it = 0
        r = None
        t = time.perf_counter()
        for _ in range(steps):
            r = p.solve(u, precision=pr, **sk)
            it += r["stats"]["iters"]
        return it, time.perf_counter() - t, r

    iters, dt, res = timed(prec, args.steps, args.warmup)
    log("timed %d solves: %.4fs, %d iterations" % (args.steps, dt, iters))
    other = H.PREC_FP32 if prec == H.PREC_FP64 else H.PREC_FP64
    it2, dt2, res2 = timed(other, max(args.steps // 2, 1), 1)

    prof = p.profile(u, precision=prec, **sk)
    kinds = [k for k in prof if k != "iters"]
    dom = max(kinds, key=lambda k: prof[k]["ms_total"])
    d = prof[dom]
    achieved = d["bytes_per_launch"] / (d["ms_avg"] * 1e-3) / 1e9
    for k in kinds:
        if prof[k]["launches"]:
            log("kernel %-16s launches=%5d avg=%.4f ms  %.1f GB/s" % (
                k, prof[k]["launches"], prof[k]["ms_avg"],
                prof[k]["bytes_per_launch"] / (prof[k]["ms_avg"] * 1e-3) / 1e9))

    out = {
        "metric": METRIC,
        "value": iters / dt,
        "unit": "EM iterations/s (SQUAREM cycles of 3 EM maps)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 E-step / f64 M-step+SQUAREM" if prec == H.PREC_FP32 else "f64",
        "other_precision": {
            "dtype": "f32 E-step / f64 M-step+SQUAREM" if other == H.PREC_FP32 else "f64",
            "value": it2 / dt2,
            "ms_per_step": dt2 / max(args.steps // 2, 1) * 1e3,
            "iters_per_solve": res2["stats"]["iters"],
        },
        "data": "synthetic",
        "config": {
            "workload": "plain SQUAREM EM, one solve per step",
            "n_tx": cls.n_tx, "n_cls_drawn": args.n_cls,
            "n_cls_multi": cls.n_cls, "nnz": cls.nnz,
            "n_frag": args.n_frag, "seed": args.seed,
            "iters_per_solve": res["stats"]["iters"],
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
            "traffic": None,
            "bytes_per_launch": d["bytes_per_launch"],
            "ms_per_launch": d["ms_avg"],
        },
        "kernels": {k: dict(prof[k]) for k in kinds if prof[k]["launches"]},
        "upload_s": t_upload,
    }

    if args.cpu_solves > 0:
        from oracle import oracle as O
        n = args.cpu_solves
        t0 = O.now()
        k_cpu = 0
        ora = None
        for _ in range(n):
            ora = O.estimate_count(cls, u, stop_rel=args.stop_rel)
            k_cpu += ora["k"]
        dt_cpu = O.now() - t0
        out["cpu_baseline"] = {
            "value": k_cpu / dt_cpu,
            "unit": "EM iterations/s",
            "cores": 1,
            "kind": "port",
            "sample": "%d full plain solves of this workload (%d iterations, "
                      "%.1fs) by the FP64 C restatement of src/EM.c, "
                      "gcc -O2, 1 thread" % (n, k_cpu, dt_cpu),
        }
        m = ora["tpm"] / 1e6 > 1e-7
        ll_o = O.loglik(cls, u, ora["lam"])
        out["parity"] = {"k_oracle": ora["k"],
                         "abund_threshold": "tpm/1e6 > 1e-7"}
        for pr, name in ((H.PREC_FP64, "fp64"), (H.PREC_FP32, "fp32")):
            g = p.solve(u, precision=pr, min_iter=ora["k"],
                        max_iter=ora["k"], **sk)
            e = np.abs(g["tpm"][m] - ora["tpm"][m]) / ora["tpm"][m]
            out["parity"][name] = {
                "k_gpu_stop_rule": (res if pr == prec else res2)["stats"]["iters"],
                "abund_rel_err_max": float(e.max()),
                "abund_rel_err_median": float(np.median(e)),
                "abund_frac_over_1e-6": float(np.mean(e > 1e-6)),
                "loglik_rel_err": abs(g["stats"]["loglik"] - ll_o) / abs(ll_o),
            }
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    p.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
