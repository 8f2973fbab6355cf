"""Serving loop: admission by HBM footprint, device-resident jobs whose work
overlaps, min-ESS target by continuing the sampling phase, and the same
result for a job alone and in a mix.

Stated tolerance for "the same result alone and in the mix": bitwise (a
job's kernels, launch shapes, Philox streams and stopping rule depend only
on the job itself)."""
import sys
import time

import numpy as np
import pytest

import stan_amd as S
from stan_amd import analyze
from stan_amd import workloads as W

pytestmark = pytest.mark.gpu

GB = 1 << 30


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _jobs():
    a = W.logistic(3000, 10, 0.5, 0.5, seed=101)
    b = W.normal(20000, 64, seed=102)
    c = W.logistic(4000, 300, 1.0 / np.sqrt(300), 0.5, seed=103)
    d = W.hier_logistic(J=20, K=3, seed=104, mean_log_size=np.log(60))
    common = dict(num_warmup=150, num_samples=100, random_seed=7)
    return {
        "a": dict(family=S.LOGIT, X=a["X"], y=a["y"], chains=8, prior=a["prior"], chain=1,
                  **common),
        "b": dict(family=S.NORMAL, X=b["X"], y=b["y"], chains=16, prior=b["prior"],
                  grad_precision=S.BF16X2, chain=1, **common),
        "c": dict(family=S.LOGIT, X=c["X"], y=c["y"], chains=4, prior=c["prior"],
                  grad_precision=S.BF16X2, chain=1, **common),  # K = 300: two-launch MFMA
        "d": dict(family=S.HIER_LOGIT, X=d["X"], y=d["y"], offsets=d["offsets"], chains=8,
                  prior=d["prior"], grad_precision=S.FP32, chain=1, **common),
    }


def _run(srv, spec):
    spec = dict(spec)
    fam, X, y, chains = spec.pop("family"), spec.pop("X"), spec.pop("y"), spec.pop("chains")
    return srv.submit(fam, X, y, chains, **spec)


def test_job_same_alone_and_in_mix():
    jobs = _jobs()
    alone = {}
    for name in ("a", "d"):
        srv = S.Server(64 * GB)
        jid = _run(srv, jobs[name])
        assert srv.wait(jid, 120)
        alone[name] = srv.results(jid)
        srv.close()
        _say(f"alone {name}: {alone[name]['info']['steps']} steps")
    srv = S.Server(64 * GB)
    ids = {name: _run(srv, jobs[name]) for name in ("b", "a", "c", "d")}
    for name, jid in ids.items():
        assert srv.wait(jid, 300), name
    infos = {name: srv.info(jid) for name, jid in ids.items()}
    # the jobs were resident at the same time
    assert max(i["t_admit"] for i in infos.values()) < min(i["t_done"] for i in infos.values())
    for name in ("a", "d"):
        mixed = srv.results(ids[name])
        assert np.all(mixed["status"] == 0)
        np.testing.assert_array_equal(mixed["rows"], alone[name]["rows"])
        np.testing.assert_array_equal(mixed["draws"], alone[name]["draws"])
    for name in ("b", "c"):
        r = srv.results(ids[name])
        assert np.all(r["status"] == 0) and np.all(r["rows"] == 100)
    # gradient precision the server chose: BF16 hi+lo where the job allows
    # it, K <= 512 and chains x K above the MFMA tile's cost (b: 16 x 64 =
    # 1024 > 0.6 x (72 + 2 x 64) for its one 64-chain tile; c: 4 x 300 =
    # 1200 > 0.6 x (72 + 2 x 320)); FP64 for jobs that ask for FP64 (a);
    # FP32 as asked for the hierarchical job (d)
    prec = {name: infos[name]["grad_precision"] for name in infos}
    assert prec == {"a": S.FP64, "b": S.BF16X2, "c": S.BF16X2, "d": S.FP32}, prec
    srv.close()


def test_extension_continues_the_same_run():
    a = W.logistic(2000, 6, 0.5, 0.5, seed=201)
    srv = S.Server(64 * GB)
    base = dict(num_warmup=120, random_seed=11, chain=1, save_warmup=1)
    j1 = srv.submit(S.LOGIT, a["X"], a["y"], 4, prior=a["prior"], num_samples=60,
                    target_min_ess=1e9, max_samples=100, **base)
    j2 = srv.submit(S.LOGIT, a["X"], a["y"], 4, prior=a["prior"], num_samples=100, **base)
    assert srv.wait(j1, 120) and srv.wait(j2, 120)
    r1, r2 = srv.results(j1), srv.results(j2)
    assert r1["info"]["extensions"] == 1 and r1["info"]["sampling_iterations"] == 100
    np.testing.assert_array_equal(r1["rows"], [220] * 4)
    np.testing.assert_array_equal(r1["draws"], r2["draws"])
    srv.close()


def test_admission_by_footprint():
    a = W.logistic(50000, 40, 0.2, 0.5, seed=301)
    spec, _ = S.job_spec(S.LOGIT, a["X"], a["y"], 8, num_warmup=100, num_samples=50)
    fp = S.job_footprint(spec)
    srv = S.Server(int(1.5 * fp))
    j1 = srv.submit(S.LOGIT, a["X"], a["y"], 8, num_warmup=100, num_samples=50, random_seed=1)
    j2 = srv.submit(S.LOGIT, a["X"], a["y"], 8, num_warmup=100, num_samples=50, random_seed=2)
    assert srv.wait(j1, 120) and srv.wait(j2, 120)
    i1, i2 = srv.info(j1), srv.info(j2)
    assert i1["footprint_bytes"] == fp
    assert i2["t_admit"] >= i1["t_done"]  # only one fits at a time
    with pytest.raises(S.StanxError):
        big = np.zeros((200000, 40))
        srv.submit(S.LOGIT, big, np.zeros(200000, dtype=np.int32), 8, num_samples=50)
    srv.close()


def test_min_ess_target_and_native_ess():
    a = W.logistic(3000, 5, 0.5, 0.5, seed=401)
    srv = S.Server(64 * GB)
    jid = srv.submit(S.LOGIT, a["X"], a["y"], 8, prior=a["prior"], num_warmup=200,
                     num_samples=30, target_min_ess=400, max_samples=2000, random_seed=3)
    assert srv.wait(jid, 120)
    r = srv.results(jid)
    inf = r["info"]
    _say(f"target job: min ESS {inf['min_ess']:.0f} after {inf['sampling_iterations']} "
         f"iterations, {inf['extensions']} extensions")
    assert inf["min_ess"] >= 400 and inf["extensions"] >= 1
    n = inf["sampling_iterations"]
    q = np.transpose(r["draws"][:, :n, 7:], (1, 0, 2))
    assert abs(analyze.min_bulk_ess(q) - inf["min_ess"]) <= 1e-8 * inf["min_ess"]
    srv.close()


def test_served_posterior_matches_oracle_sampler():
    """A job run through the serving loop, sharing the device with another
    job, against the oracle sampler on the same posterior: means and sds
    within 4 MCSE."""
    a = W.logistic(2000, 5, 1.0 / np.sqrt(5), 0.5, seed=501)
    other = W.normal(20000, 32, seed=502)
    srv = S.Server(64 * GB)
    jid = srv.submit(S.LOGIT, a["X"], a["y"], 8, prior=a["prior"], num_warmup=500,
                     num_samples=1000, random_seed=55, chain=1)
    oid = srv.submit(S.NORMAL, other["X"], other["y"], 16, prior=other["prior"],
                     grad_precision=S.BF16X2, num_warmup=200, num_samples=200, random_seed=56)
    assert srv.wait(jid, 300) and srv.wait(oid, 300)
    r = srv.results(jid)
    srv.close()
    assert np.all(r["status"] == 0)
    dev = np.transpose(r["draws"][:, :1000, 7:], (1, 0, 2))
    import oracle as O
    ref = O.GlmData(O.LOGIT, a["X"], a["y"], prior=a["prior"])
    f = O.model_fn("orc_glm_model_cb")
    ctx = ref.ctx(mode=0)
    cpu = []
    for chain in range(1, 5):
        rc, dr, _, _, _ = O.hmc_nuts_diag_e_adapt(f, ctx, 6, seed=5555, chain=chain,
                                                  num_warmup=500, num_samples=1000)
        assert rc == 0
        cpu.append(dr[:, 7:])
    cpu = np.stack(cpu, 1)
    for p in range(6):
        se = np.hypot(analyze.mcse_mean(dev[:, :, p]), analyze.mcse_mean(cpu[:, :, p]))
        assert abs(dev[:, :, p].mean() - cpu[:, :, p].mean()) < 4 * se, p
        se_sd = np.hypot(analyze.mcse_sd(dev[:, :, p]), analyze.mcse_sd(cpu[:, :, p]))
        assert abs(dev[:, :, p].std() - cpu[:, :, p].std()) < 4 * se_sd, p
    _say(f"served job: min bulk ESS {analyze.min_bulk_ess(dev):.0f}")
