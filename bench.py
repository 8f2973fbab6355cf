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


def _nanmax(a, b):
    """max that keeps a NaN (the builtin max(a, nan) returns a)."""
    return a if a != a else (b if b != b else max(a, b))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _host_threads():
    """Threads the CPU leg may use: the cores this process may run on,
    bounded by OMP_NUM_THREADS when it is set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(d, theta, budget_s=10.0):
    """The CPU restatement of the reference's two-pass GLM evaluation
    (oracle/glm_cpu_baseline.c: -O3, x86-64-v3, OpenMP over rows), timed on
    a bounded number of evaluations of the same data, on every host core
    this process may use and on one core."""
    import ctypes as C
    lib = C.CDLL(os.path.join(ROOT, "oracle", "libcpu_baseline.so"))
    f = lib.cpu_logit_lpgrad
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_double,
                  C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_void_p]
    X = np.ascontiguousarray(d["X"])
    y = np.ascontiguousarray(d["y"], dtype=np.int32)
    N, K = X.shape
    th = np.ascontiguousarray(theta, dtype=np.float64)
    g = np.zeros(K + 1)
    lp = C.c_double()

    def rate(threads, budget):
        f(X.ctypes.data, y.ctypes.data, N, K, d["prior"][0], d["prior"][1], th.ctypes.data,
          threads, C.byref(lp), g.ctypes.data)
        n, t0 = 0, time.perf_counter()
        while True:
            f(X.ctypes.data, y.ctypes.data, N, K, d["prior"][0], d["prior"][1], th.ctypes.data,
              threads, C.byref(lp), g.ctypes.data)
            n += 1
            el = time.perf_counter() - t0
            if el > budget or n >= 500:
                return n / el, n, el
    cores = _host_threads()
    v_all, n_all, el_all = rate(cores, budget_s)
    v_one, n_one, el_one = rate(1, budget_s / 2)
    return dict(value=v_all, unit="grad evals/s", cores=cores, kind="port",
                sample=f"{n_all} evaluations of the N=1e6 K=100 logistic log density + gradient "
                       f"(two-pass FP64 CPU restatement, -O3 x86-64-v3, OpenMP over rows) on "
                       f"{cores} threads in {el_all:.1f} s",
                one_core=dict(value=v_one, cores=1,
                              sample=f"{n_one} evaluations on one thread in {el_one:.1f} s"))


def nuts_run(model, warmup, samples, seed=1234, **kw):
    import stan_amd as S
    from stan_amd import analyze
    out = S.hmc_nuts_diag_e_adapt(model, random_seed=seed, chain=1, num_warmup=warmup,
                                  num_samples=samples, **kw)
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


def batched_run(model, chains, warmup, samples, grad_precision, seed=1234, **kw):
    """Device-resident NUTS over `chains` chains; min-ESS over pooled chains.
    Runs in bounded pieces so that progress is reported while it runs."""
    import stan_amd as S
    from stan_amd import analyze
    b = S.NutsBatch(model, chains, grad_precision=grad_precision, random_seed=seed, chain=1,
                    num_warmup=warmup, num_samples=samples, **kw)
    t0 = time.perf_counter()
    last = t0
    while True:
        steps, done = b.run(max_steps=2048)
        if done >= chains or steps == 0:
            break
        if time.perf_counter() - last > 10:
            last = time.perf_counter()
            log(f"  ... {done}/{chains} chains done, {last - t0:.0f} s")
    res = b.results()
    b.close()
    if not (res["status"] == 0).all():
        raise RuntimeError(f"chain status {res['status']}")
    q = np.transpose(res["draws"][:, :, 7:], (1, 0, 2))  # iterations x chains x params
    ess = analyze.min_bulk_ess(q)
    secs = res["seconds"]
    return dict(chains=chains, min_ess=ess, seconds=secs, min_ess_per_s=ess / secs,
                grad_evals=int(res["n_grad"].sum()),
                grad_evals_per_s=float(res["n_grad"].sum()) / secs, steps=int(res["steps"]),
                ms_per_step=1e3 * secs / max(1, res["steps"]),
                mean_n_leapfrog=float(res["draws"][:, :, 4].mean()),
                mean_treedepth=float(res["draws"][:, :, 3].mean()),
                stepsize_median=float(np.median(res["stepsize"])), warmup=warmup,
                samples=samples)


def serving_stream(window_s, rate=20.0, seed=7):
    """Poisson arrivals (rate jobs/s) over the first window_s seconds of the
    stream; per job: logistic or normal GLM with equal probability, N
    log-uniform in [1e3, 2e6], K log-uniform in [8, 512], chains uniform in
    {4, 8, 16, 32, 64}.  The draws are those of the stream's first jobs
    whatever the window (one generator call per quantity, 4096 jobs)."""
    n_max = 4096
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.cumsum(rng.exponential(1.0 / rate, n_max))
    logit = rng.random(n_max) < 0.5
    N = np.exp(rng.uniform(np.log(1e3), np.log(2e6), n_max)).astype(np.int64)
    K = np.exp(rng.uniform(np.log(8), np.log(512), n_max)).astype(np.int64)
    C = rng.choice([4, 8, 16, 32, 64], n_max)
    n = int(np.searchsorted(t, window_s))
    return [dict(t=float(t[i]), logit=bool(logit[i]), N=int(N[i]), K=int(K[i]), C=int(C[i]),
                 seed=1000 + i) for i in range(n)]


def serving_data(job):
    """X ~ N(0, 1/K); logistic: beta ~ N(0,1), alpha = 0.5; normal: sigma = 1."""
    rng = np.random.Generator(np.random.PCG64(job["seed"]))
    N, K = job["N"], job["K"]
    X = rng.standard_normal((N, K), dtype=np.float32).astype(np.float64) / np.sqrt(K)
    beta = rng.standard_normal(K)
    eta = X @ beta + 0.5
    if job["logit"]:
        y = (rng.random(N) < 1.0 / (1.0 + np.exp(-eta))).astype(np.int32)
    else:
        y = eta + rng.standard_normal(N)
    return X, y


def serving_section(window_s=3.0, target_ess=1000.0, warmup=200, deadline_s=60.0):
    """C5: the arrival stream at 20 jobs/s for window_s seconds, replayed in
    real time against the serving loop (jobs queue while the device is
    busy), with a deadline of deadline_s after the last arrival.  Latency is
    submit to result and includes the queueing; jobs unfinished at the
    deadline are reported, and then the p99 is only a lower bound.  The
    sustainable rate is completed jobs per second of the busy period."""
    import concurrent.futures as cf
    import stan_amd as S
    from stan_amd import workloads as W
    jobs = serving_stream(window_s)
    log(f"serving: generating the {len(jobs)} jobs of a {window_s:.0f} s window")
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        data = list(ex.map(serving_data, jobs))
    srv = S.Server(200 * (1 << 30), max_active=16)
    ids = []
    t0 = time.perf_counter()
    for j in jobs:
        delay = j["t"] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        X, y = data[len(ids)]
        fam = S.LOGIT if j["logit"] else S.NORMAL
        prior = W.PRIOR_LOGISTIC if j["logit"] else W.PRIOR_NORMAL
        first = int(np.ceil(1.5 * target_ess / j["C"])) + 20
        ids.append(srv.submit(fam, X, y, j["C"], prior=prior, grad_precision=S.BF16X2,
                              target_min_ess=target_ess, max_samples=max(400, 8 * first),
                              num_warmup=warmup, num_samples=first, random_seed=j["seed"],
                              chain=1))
        data[len(ids) - 1] = None  # the server holds its own copy until admission
    t_submitted = time.perf_counter()
    lag = t_submitted - t0 - jobs[-1]["t"]
    log(f"serving: {len(ids)} jobs submitted over {t_submitted - t0:.1f} s "
        f"(submission lag {lag:.2f} s)")
    last = t_submitted
    pending = list(ids)
    while pending and time.perf_counter() - t_submitted < deadline_s:
        pending = [i for i in pending if not srv.wait(i, 0.0)]
        time.sleep(0.2)
        if time.perf_counter() - last > 10:
            last = time.perf_counter()
            log(f"  ... serving: {len(ids) - len(pending)}/{len(ids)} jobs done")
    infos_all = {i: srv.info(i) for i in ids}
    srv.close()  # cancels what is still running at the deadline
    unfinished = len(pending)
    infos = [infos_all[i] for i in ids if i not in pending]
    out = dict(window_s=window_s, jobs=len(ids), completed=len(infos),
               unfinished_at_deadline=unfinished, deadline_s=deadline_s,
               arrival_rate_per_s=20.0, seed=7, target_min_ess=target_ess, warmup=warmup)
    if not infos:
        return out
    lat = np.array([i["t_done"] - i["t_submit"] for i in infos])
    # unfinished jobs count as latencies beyond the deadline
    cens = np.concatenate([lat, np.full(unfinished, np.inf)])
    queue = np.array([i["t_admit"] - i["t_submit"] for i in infos])
    t_first = min(i["t_submit"] for i in infos_all.values())
    busy = max(i["t_done"] for i in infos) - t_first
    n_grad = sum(i["n_grad"] for i in infos)
    ess = [i["min_ess"] for i in infos]
    out.update(p50_s=float(np.percentile(cens, 50)), p99_s=float(np.percentile(cens, 99)),
               p99_is_lower_bound=bool(unfinished > 0), max_completed_s=float(lat.max()),
               p50_queue_s=float(np.percentile(queue, 50)),
               p99_queue_s=float(np.percentile(queue, 99)),
               busy_period_s=busy, sustainable_jobs_per_s=len(infos) / busy,
               grad_evals_per_s=n_grad / busy, min_ess_per_s=float(np.nansum(ess)) / busy,
               jobs_reaching_target=int(sum(e >= target_ess for e in ess)),
               min_ess_min=float(np.nanmin(ess)),
               extensions=int(sum(i["extensions"] for i in infos)),
               mfma_jobs=int(sum(1 for i in infos if i["grad_precision"] == S.BF16X2)))
    return out


def mixed_error(model, ref_data, family, thetas):
    """Max relative lp error and normwise gradient error of the batched
    mixed-precision pass against the FP64 oracle at the given points."""
    import oracle as O
    ref = O.GlmData(family, ref_data["X"], ref_data["y"], offsets=ref_data.get("offsets"),
                    prior=ref_data["prior"])
    lp, g = model.log_prob_grad_batch(thetas)
    e_lp = e_g = 0.0
    for c in range(len(thetas)):
        st, lp_r, g_r = ref.lpgrad(thetas[c], mode=1)
        e_lp = _nanmax(e_lp, abs(lp[c] - lp_r) / abs(lp_r))
        e_g = _nanmax(e_g, float(np.linalg.norm(g[c] - g_r) / np.linalg.norm(g_r)))
    return e_lp, e_g


def split_rhat(draws):
    """draws: iterations x chains x params -> split R-hat per parameter."""
    n = draws.shape[0] // 2
    s = np.concatenate([draws[:n], draws[n:2 * n]], axis=1)
    w = s.var(axis=0, ddof=1).mean(axis=0)
    b = n * s.mean(axis=0).var(axis=0, ddof=1)
    return np.sqrt(((n - 1) / n * w + b / n) / w)


def hier_section(warm, samp, chains=64, max_depth=10, plain_samples=200):
    """Hierarchical logistic model, J=1000 ragged groups, K=20: batched FP32
    evaluation and device NUTS over 64 chains, against the single-chain FP64
    service on the same context (same warm-up, fewer draws)."""
    import stan_amd as S
    from stan_amd import workloads as W
    dh = W.hier_logistic()
    N, K = dh["X"].shape
    J = len(dh["offsets"]) - 1
    log(f"hierarchical logistic J={J} N={N} K={K}, {chains} chains FP32, "
        f"{warm} + {samp} iterations, max_depth {max_depth}")
    mh = S.GlmModel(S.HIER_LOGIT, dh["X"], dh["y"], offsets=dh["offsets"], prior=dh["prior"],
                    precision=S.FP32)
    th = W.random_thetas(mh.D, chains, seed=9) * 0.5
    ms_b, _ = mh.time_log_prob_grad(th, iters=20, warmup=3, chains=chains)
    ms_1, _ = mh.time_log_prob_grad(th[0], iters=20, warmup=3)
    b = S.NutsBatch(mh, chains, grad_precision=S.FP32, random_seed=1234, chain=1,
                    num_warmup=warm, num_samples=samp, max_depth=max_depth)
    t0 = last = time.perf_counter()
    while True:
        steps, done = b.run(max_steps=2048)
        if done >= chains or steps == 0:
            break
        if time.perf_counter() - last > 10:
            last = time.perf_counter()
            log(f"  ... {done}/{chains} chains done, {last - t0:.0f} s")
    res = b.results()
    b.close()
    if not (res["status"] == 0).all():
        raise RuntimeError(f"chain status {res['status']}")
    from stan_amd import analyze
    q = np.transpose(res["draws"][:, :, 7:], (1, 0, 2))
    ess = analyze.min_bulk_ess(q)
    secs = res["seconds"]
    rh = split_rhat(q)
    out = dict(chains=chains, min_ess=ess, seconds=secs, min_ess_per_s=ess / secs,
               grad_evals=int(res["n_grad"].sum()),
               grad_evals_per_s=float(res["n_grad"].sum()) / secs, steps=int(res["steps"]),
               ms_per_step=1e3 * secs / max(1, res["steps"]),
               mean_n_leapfrog=float(res["draws"][:, :, 4].mean()),
               mean_treedepth=float(res["draws"][:, :, 3].mean()),
               stepsize_median=float(np.median(res["stepsize"])),
               stepsize_max_over_median=float(np.max(res["stepsize"]) /
                                               np.median(res["stepsize"])),
               divergent_draws=int(res["draws"][:, :, 5].sum()),
               split_rhat_max=float(rh.max()),
               split_rhat_max_global=float(rh[: 2 * K + 2].max()),
               warmup=warm, samples=samp)
    # X (FP32) + y (int8) + chunk tables; theta/out of every chain
    alg_bytes = 4.0 * N * K + N + 16.0 * (N / 256 + J) + 8.0 * chains * (2 * mh.D + 2)
    out.update(N=N, K=K, J=J, D=mh.D, max_depth=max_depth, ms_per_batched_eval=ms_b,
               batched_eval_gbs=alg_bytes / (ms_b * 1e-3) / 1e9,
               batched_eval_tflops_fp32=(4.0 * N * K + 4.0 * N) * chains / (ms_b * 1e-3) / 1e12,
               ms_single_fp64_eval=ms_1)
    e_lp, e_g = mixed_error(mh, dh, S.HIER_LOGIT, th[:3])
    out["max_rel_err_lp"] = e_lp
    out["max_rel_err_grad"] = e_g
    log(f"hierarchical: single-chain FP64 service for comparison ({warm} + {plain_samples})")
    out["plain_fp64_single_chain"] = nuts_run(mh, warm, plain_samples, max_depth=max_depth)
    out["min_ess_per_s_vs_plain"] = (out["min_ess_per_s"] /
                                     out["plain_fp64_single_chain"]["min_ess_per_s"])
    mh.close()
    return out


def c1_section(with_cpu=True):
    """The reference's own CPU-runnable case: logistic N=1e4, K=50, X~N(0,1),
    4 chains x (1000 warmup + 1000 draws), max_depth 10, seed 1234.  The
    device sampler runs the 4 chains together with FP64 gradients; the
    oracle sampler runs them on 4 host threads."""
    import stan_amd as S
    from stan_amd import analyze
    from stan_amd import workloads as W
    d = W.logistic_small()
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    out = dict(gpu_batched_fp64=batched_run(m, 4, 1000, 1000, S.FP64))
    m.close()
    if with_cpu:
        import threading
        import oracle as O
        ref = O.GlmData(O.LOGIT, d["X"], d["y"], prior=d["prior"])
        f = O.model_fn("orc_glm_model_cb")
        ctx = ref.ctx(mode=0)
        res = [None] * 4

        def run(c):
            res[c] = O.hmc_nuts_diag_e_adapt(f, ctx, 51, seed=1234, chain=c + 1,
                                             num_warmup=1000, num_samples=1000)
        log("C1 on the host: oracle sampler, 4 chains on 4 threads")
        t0 = time.perf_counter()
        ths = [threading.Thread(target=run, args=(c,)) for c in range(4)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        el = time.perf_counter() - t0
        if any(r[0] != 0 for r in res):
            raise RuntimeError("oracle sampler failed")
        q = np.stack([r[1][:, 7:] for r in res], 1)
        ess = analyze.min_bulk_ess(q)
        ng = sum(r[4] for r in res)
        out["cpu_oracle_4_threads"] = dict(seconds=el, min_ess=ess, min_ess_per_s=ess / el,
                                           grad_evals=ng, grad_evals_per_s=ng / el, cores=4,
                                           kind="port")
        out["gpu_vs_cpu_min_ess_per_s"] = (out["gpu_batched_fp64"]["min_ess_per_s"] /
                                           out["cpu_oracle_4_threads"]["min_ess_per_s"])
    return out


def kernels_section(steps, sampler_iters=30, only=None):
    """Short runs of the batched kernels alone, sized for a kernel trace:
    `steps` batched evaluations of each batched workload (logistic N=1e6
    K=100 with 256 chains and normal N=2e5 K=256 with 256 chains on BF16
    hi+lo MFMA; hierarchical logistic J=1000 K=20 with 64 chains in FP32),
    then `sampler_iters` warm-up + `sampler_iters` draws of the device
    sampler on each.  Prints one JSON line of milliseconds per batched
    evaluation; no CPU leg."""
    import stan_amd as S
    from stan_amd import workloads as W
    out = {}
    cases = [("logistic_n1e6_k100_c256", S.LOGIT, W.logistic_large, 256, S.BF16X2),
             ("normal_n2e5_k256_c256", S.NORMAL, W.normal_batched, 256, S.BF16X2),
             ("hier_logistic_j1000_k20_c64", S.HIER_LOGIT, W.hier_logistic, 64, S.FP32)]
    for name, fam, gen, C, prec in cases:
        if only and name not in only:
            continue
        log(f"kernels: {name}")
        d = gen()
        m = S.GlmModel(fam, d["X"], d["y"], offsets=d.get("offsets"), prior=d["prior"],
                       precision=prec)
        for line in m.tuning_log().splitlines():
            if "chosen" in line or "cached" in line:
                log(f"  {line}")
        th = W.random_thetas(m.D, C, seed=7) * 0.5
        ms, _ = m.time_log_prob_grad(th, iters=steps, warmup=3, chains=C)
        out[name] = dict(ms_per_batched_eval=ms, chains=C)
        if sampler_iters > 0:
            r = batched_run(m, C, sampler_iters, sampler_iters, prec, max_depth=6)
            out[name].update(sampler_steps=r["steps"], sampler_ms_per_step=r["ms_per_step"])
        m.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-nuts", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--chains", type=int, default=256)
    ap.add_argument("--no-hier", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-window", type=float, default=3.0,
                    help="seconds of the 20 jobs/s arrival stream to replay")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="only the copy kernel of known size (the measured HBM ceiling; run "
                         "under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE it gives the counter "
                         "scale)")
    ap.add_argument("--kernels", action="store_true",
                    help="only short runs of the batched kernels (for kernel traces)")
    ap.add_argument("--case", action="append",
                    help="with --kernels: run only the named workload (repeatable)")
    ap.add_argument("--sampler-iters", type=int, default=30,
                    help="with --kernels: warm-up and draws of the device sampler per case "
                         "(0: evaluations only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(across GPUs, the plain path would shard rows and the batched "
                          "sampler whole chains)"}))
        sys.exit(2)
    if args.calibrate:
        import stan_amd as S
        nbytes = 800_000_000
        gbs, ms = S.device_copy_bandwidth(nbytes, iters=max(3, args.steps))
        print(json.dumps({"calibrate": dict(copy_bytes=nbytes, read_kib_per_copy=nbytes / 1024,
                                            write_kib_per_copy=nbytes / 1024, copies=max(3, args.steps) + 1,
                                            ms_per_copy=ms, gbs_read_plus_write=gbs,
                                            peak_gbs=PEAK_HBM_GBS)}))
        return
    if args.kernels:
        print(json.dumps({"kernels": kernels_section(args.steps, args.sampler_iters,
                                                     args.case)}))
        return

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
    copy_gbs, _ = S.device_copy_bandwidth(800_000_000, iters=20)
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=PEAK_HBM_GBS,
                    unit="GB/s", frac=round(achieved / PEAK_HBM_GBS, 4), traffic=traffic,
                    kernel="glm_pass_f64 (logistic)", bytes_per_launch=alg_bytes,
                    kernel_ms=ms_kernel, kernel_config=m.kernel_config(),
                    measured_copy_gbs=round(copy_gbs, 1),
                    frac_of_measured_copy=round(achieved / copy_gbs, 4))

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
    warm, samp = 200, 300
    if not args.no_nuts:
        log(f"plain NUTS run (1 chain, FP64, {warm} warmup + {samp} samples)")
        result["nuts_plain_fp64"] = nuts_run(m, warm, samp)
    m.close()
    if not args.no_batched:
        log(f"batched mixed-precision NUTS on the same data ({args.chains} chains)")
        mb = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"], precision=S.BF16X2)
        C = args.chains
        th = W.random_thetas(mb.D, C, seed=7)
        ms_b, _ = mb.time_log_prob_grad(th, iters=20, warmup=3, chains=C)
        bat = batched_run(mb, C, warm, samp, S.BF16X2)
        bat["ms_per_batched_eval"] = ms_b
        Kp = (K + 31) // 32 * 32
        bat["mfma_tflops_algorithmic"] = 4.0 * N * K * C / (ms_b * 1e-3) / 1e12
        bat["mfma_tflops_issued"] = 3 * 4.0 * N * Kp * C / (ms_b * 1e-3) / 1e12
        if not args.no_nuts:
            bat["min_ess_per_s_vs_plain"] = (bat["min_ess_per_s"] /
                                            result["nuts_plain_fp64"]["min_ess_per_s"])
        log("mixed-precision error vs FP64 oracle at 3 points")
        e_lp, e_g = mixed_error(mb, d, S.LOGIT, th[:3] * 0.5)
        bat["max_rel_err_lp"] = e_lp
        bat["max_rel_err_grad"] = e_g
        result["nuts_batched_mixed"] = bat
        mb.close()

        log("normal GLM, N=2e5, K=256, 256 chains (500 warmup + 500 samples)")
        dn = W.normal_batched()
        mn = S.GlmModel(S.NORMAL, dn["X"], dn["y"], prior=dn["prior"], precision=S.BF16X2)
        thn = W.random_thetas(mn.D, 256, seed=8) * 0.5
        ms_n, _ = mn.time_log_prob_grad(thn, iters=20, warmup=3, chains=256)
        bn = batched_run(mn, 256, 500, 500, S.BF16X2)
        Nn, Kn = dn["X"].shape
        bn["ms_per_batched_eval"] = ms_n
        bn["mfma_tflops_algorithmic"] = 4.0 * Nn * Kn * 256 / (ms_n * 1e-3) / 1e12
        bn["mfma_frac_bf16_peak"] = 3 * bn["mfma_tflops_algorithmic"] / 2500.0
        e_lp, e_g = mixed_error(mn, dn, S.NORMAL, thn[:4])
        bn["max_rel_err_lp"] = e_lp
        bn["max_rel_err_grad"] = e_g
        result["normal_glm_batched_256"] = bn
        mn.close()
    if not args.no_serve:
        log(f"serving: {args.serve_window:.0f} s of the 20 jobs/s arrival stream")
        result["serving"] = serving_section(args.serve_window)
    if not args.no_hier:
        result["hier_logistic_64"] = hier_section(1000, 500)
    if not args.no_c1:
        log("reference CPU case N=1e4 K=50, 4 chains x (1000 + 1000)")
        result["c1_logistic_4_chains"] = c1_section(with_cpu=not args.no_cpu)
    if not args.no_cpu:
        log("cpu baseline (oracle, 1 core)")
        result["cpu_baseline"] = cpu_baseline(d, theta)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
