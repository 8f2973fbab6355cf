"""Benchmark: gradient evaluations per second of the GLM log density on one
MI355X, plus min-ESS/s of the NUTS sampler that drives it.

Workload (default): Bayesian logistic regression, N=1,000,000 rows, K=100
features, X ~ N(0, 1/K), FP64, one chain -- the plain drop-in path.  A
"step" is one log_prob + gradient evaluation over the whole data set with
X already resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the bulk logistic pass at N=1e6, K=100 with the
# lanes_per_row=4 layout, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes,
# each scaled by the same counter on a plain copy of known size (fetch x2,
# write x1).  Summary and commands: profiles/glm_pass_f64_logistic_n1e6_k100_pmc.md
PMC_TRAFFIC_BYTES = {(4, 25): 783350 * 1024.0 + 3298 * 1024.0}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(d, theta, budget_s=12.0):
    """The oracle's Stan-equivalent two-pass FP64 evaluation (mode 0), one
    host core, timed on a bounded number of evaluations of the same data."""
    import oracle as O
    ref = O.GlmData(O.LOGIT, d["X"], d["y"], prior=d["prior"])
    ref.lpgrad(theta, mode=0)
    n, t0 = 0, time.perf_counter()
    while True:
        ref.lpgrad(theta, mode=0)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 200:
            break
    return dict(value=n / el, unit="grad evals/s", cores=1, kind="port",
                sample=f"{n} evaluations of the N=1e6 K=100 logistic log density + gradient "
                       f"(oracle two-pass FP64, one thread), {el:.1f} s")


def nuts_run(model, warmup, samples, seed=1234):
    import stan_amd as S
    from stan_amd import analyze
    out = S.hmc_nuts_diag_e_adapt(model, random_seed=seed, chain=1, num_warmup=warmup,
                                  num_samples=samples)
    if out["rc"] != 0:
        raise RuntimeError(out["error"])
    q = out["draws"][:, 7:]
    ess = analyze.min_bulk_ess(q[:, None, :])
    total = out["warmup_seconds"] + out["sampling_seconds"]
    return dict(min_ess=ess, seconds=total, sampling_seconds=out["sampling_seconds"],
                min_ess_per_s=ess / total, n_grad=out["n_grad"],
                grad_evals_per_s=out["n_grad"] / total, stepsize=out["stepsize"],
                mean_n_leapfrog=float(out["draws"][:, 4].mean()),
                warmup=warmup, samples=samples)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-nuts", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(the multi-GPU design shards rows or chains, see DESIGN.md)"}))
        sys.exit(2)

    import stan_amd as S
    from stan_amd import workloads as W

    log("generating logistic N=1e6 K=100")
    d = W.logistic_large()
    N, K = d["X"].shape
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    theta = W.random_thetas(m.D, 1, seed=2024)[0]
    log(f"kernel config {m.kernel_config()}")
    log("autotune:\n" + m.tuning_log())

    # warmup, then exactly `steps` timed evaluations bracketed by HIP events
    # on the context's stream (the call synchronizes before and after);
    # the same call then times `steps` launches of the bulk X pass alone
    m.time_log_prob_grad(theta, iters=max(1, args.warmup), warmup=0)
    ms_per_step, ms_kernel = m.time_log_prob_grad(theta, iters=args.steps, warmup=0)
    value = 1e3 / ms_per_step
    log(f"ms/eval {ms_per_step:.4f} (bulk X pass {ms_kernel:.4f})")

    # end-to-end host API call (PCIe round trip of theta and gradient)
    for _ in range(5):
        m.log_prob_grad(theta)
    t1 = time.perf_counter()
    nh = 50
    for _ in range(nh):
        m.log_prob_grad(theta)
    ms_host = (time.perf_counter() - t1) * 1e3 / nh

    alg_bytes = 8.0 * N * K + 1.0 * N + 8.0 * m.D
    achieved = alg_bytes / (ms_kernel * 1e-3) / 1e9
    kc = m.kernel_config()
    traffic = PMC_TRAFFIC_BYTES.get((kc["lanes_per_row"], kc["slots_per_lane"]))
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS,
                    unit="GB/s", frac=round(achieved / PEAK_HBM_GBS, 4), traffic=traffic,
                    kernel="glm_pass_f64 (logistic)", bytes_per_launch=alg_bytes,
                    kernel_ms=ms_kernel, kernel_config=m.kernel_config())

    result = {
        "metric": "grad evals/s + min-ESS/s, batched vs plain on 1 GPU; p50/p99 request latency",
        "value": round(value, 2),
        "unit": "grad evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain logistic GLM log_prob_grad, N=1e6, K=100, "
                               "X~N(0,1/K) FP64 resident in HBM, 1 chain",
                   "N": N, "K": K, "chains": 1},
        "roofline": roofline,
        "host_api_ms_per_eval": round(ms_host, 4),
    }
    if not args.no_nuts:
        log("plain NUTS run (1 chain, 200 warmup + 300 samples)")
        result["nuts_plain_fp64"] = nuts_run(m, 200, 300)
    if not args.no_cpu:
        log("cpu baseline (oracle, 1 core)")
        result["cpu_baseline"] = cpu_baseline(d, theta)
    m.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
