"""Pins the CPU oracle before it is trusted as the checker.

Every oracle function is checked against the reference's own goldens where
they exist (restated in tests/golden/reference_goldens.json) and otherwise
against closed forms, invariants, central finite differences (the
reference's recipe, src/stan/model/finite_diff_grad.hpp:130-151) and an
independent reverse-mode tape on tiny inputs.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from stan_amd import analyze
from tests import _tape as T


def test_philox_known_answers(goldens):
    for case in goldens["philox4x32_10_kat"]["cases"]:
        assert O.philox(case["ctr"], case["key"]) == case["out"]


def test_rng_stream_properties():
    u = O.rng_stream(7, 3, "u" * 20000)
    z = O.rng_stream(7, 3, "n" * 20000)
    assert np.all((u >= 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 4 * math.sqrt(1 / 12 / 20000)
    assert abs(z.mean()) < 4 / math.sqrt(20000)
    assert abs(z.var() - 1) < 4 * math.sqrt(2 / 20000)
    # different chains give different streams
    assert not np.array_equal(O.rng_stream(7, 4, "u" * 10), u[:10])


def _point(q, p, minv, V=0.0):
    arrs = [np.array(q, dtype=np.float64), np.array(p, dtype=np.float64),
            np.zeros(len(q)), np.array(minv, dtype=np.float64)]
    pt = O.Point(len(q), arrs[0].ctypes.data, arrs[1].ctypes.data, arrs[2].ctypes.data,
                 arrs[3].ctypes.data, V)
    pt._keep = arrs
    return pt, arrs


def _gauss_init(pt, arrs):
    q = arrs[0]
    pt.V = 0.5 * float(q @ q)
    arrs[2][:] = q


def test_leapfrog_goldens(goldens):
    f = O.model_fn("orc_gauss_model_cb")
    for case in goldens["expl_leapfrog"]["evolve"]:
        pt, arrs = _point([case["in"]["q"]], [case["in"]["p"]], [case["minv"]],
                          case["in"]["V"])
        arrs[2][0] = case["in"]["q"]  # g = q for the command model
        assert O.lib().orc_leapfrog(f, None, C.byref(pt), case["eps"]) == 0
        assert abs(pt.V - case["out"]["V"]) < case["tolV"], case["name"]
        assert abs(arrs[0][0] - case["out"]["q"]) < 5e-14, case["name"]
        assert abs(arrs[1][0] - case["out"]["p"]) < 5e-14, case["name"]
        assert abs(arrs[2][0] - case["out"]["q"]) < 5e-14, case["name"]


def _build_tree(depth, sign, eps, H0, q0, p0, seed=4839294):
    f = O.model_fn("orc_gauss_model_cb")
    D = len(q0)
    z, zarr = _point(q0, p0, np.ones(D))
    _gauss_init(z, zarr)
    zp, parr = _point(np.zeros(D), np.zeros(D), np.ones(D))
    psb, pse, pb, pe = (np.zeros(D) for _ in range(4))
    rho = np.array(p0, dtype=np.float64)
    nl, lsw, smp, div = C.c_int(0), C.c_double(-np.inf), C.c_double(0), C.c_int(0)
    r = O.Rng()
    O.lib().orc_rng_init(C.byref(r), seed, 0)
    v = O.lib().orc_build_tree(f, None, C.byref(r), eps, 1000.0, depth, C.byref(z),
                               C.byref(zp), psb.ctypes.data, pse.ctypes.data,
                               rho.ctypes.data, pb.ctypes.data, pe.ctypes.data, H0, sign,
                               C.byref(nl), C.byref(lsw), C.byref(smp), C.byref(div))
    return dict(valid=v, rho=rho, p_beg=pb, p_end=pe, p_sharp_beg=psb, p_sharp_end=pse,
                q=zarr[0], p=zarr[1], n_leapfrog=nl.value, lsw=lsw.value,
                sum_metro=smp.value)


def test_build_tree_golden(goldens):
    g = goldens["build_tree_gauss3d"]
    out = _build_tree(g["depth"], 1.0, g["eps"], g["H0"], g["q0"], g["p0"])
    assert out["valid"] == 1
    # EXPECT_FLOAT_EQ: 4 ulps in single precision
    for key, ref in [("rho", g["rho"]), ("p_beg", g["p_begin"]), ("p_sharp_beg", g["p_begin"]),
                     ("p_end", g["p_end"]), ("p_sharp_end", g["p_end"]), ("q", g["q_end"]),
                     ("p", g["p_end"])]:
        np.testing.assert_allclose(out[key], ref, rtol=5e-7, err_msg=key)
    assert out["n_leapfrog"] == g["n_leapfrog"]
    assert out["sum_metro"] == pytest.approx(g["sum_metro_prob"], rel=5e-7)
    assert out["lsw"] == pytest.approx(math.log(g["sum_metro_prob"]), rel=5e-7)


@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("depth", [0, 1, 2, 3])
def test_tree_boundaries_are_leapfrog_steps(depth, sign):
    """Restates tree_boundary_test (unit_e_nuts_test.cpp:97-309): the tree's
    first and last momenta are those of 1 and 2^depth explicit steps."""
    q0, p0, eps = [1.0, -1.0, 1.0], [-1.0, 1.0, -1.0], 0.1
    out = _build_tree(depth, sign, eps, -0.1, q0, p0)
    f = O.model_fn("orc_gauss_model_cb")
    z, arrs = _point(q0, p0, np.ones(3))
    _gauss_init(z, arrs)
    traj = []
    for _ in range(2 ** depth):
        O.lib().orc_leapfrog(f, None, C.byref(z), sign * eps)
        traj.append(arrs[1].copy())
    np.testing.assert_allclose(out["p_beg"], traj[0], rtol=1e-14)
    np.testing.assert_allclose(out["p_end"], traj[-1], rtol=1e-14)
    assert out["n_leapfrog"] == 2 ** depth


def test_funnel_potential(goldens):
    g = goldens["funnel_potential"]
    f = O.lib().orc_funnel_model_cb
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_void_p]
    q = np.ones(11)
    grad = np.zeros(11)
    lp = C.c_double()
    f(None, q.ctypes.data, 11, C.byref(lp), grad.ctypes.data)
    assert -lp.value == pytest.approx(g["V"], rel=5e-7)
    assert -grad[0] == pytest.approx(g["g0"], rel=5e-7)
    np.testing.assert_allclose(-grad[1:], g["gi"], rtol=5e-7)


def test_learn_stepsize_goldens(goldens):
    g = goldens["learn_stepsize"]
    fp = g["fixed_point"]
    state = np.array([0, 0, 0, fp["mu"], fp["delta"], 0.05, 0.75, 10.0])
    eps = C.c_double(0.0)
    O.lib().orc_learn_stepsize(state.ctypes.data, C.byref(eps), fp["stat"])
    assert eps.value == fp["eps_out"]
    for case in g["cases"]:
        state = np.array(case["state"], dtype=np.float64)
        eps = C.c_double(case["eps"])
        O.lib().orc_learn_stepsize(state.ctypes.data, C.byref(eps), case["stat"])
        assert abs(eps.value - case["eps_out"]) < case["tol"]
        if "counter" in case:
            assert state[0] == case["counter"]
            assert abs(state[1] - case["s_bar"]) < 1e-14
            assert abs(state[2] - case["x_bar"]) < 1e-14


# ---------------- GLM oracle ----------------

def _rand_logit(N, K, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, K)) / math.sqrt(K)
    y = (rng.random(N) < 0.4).astype(np.int32)
    return X, y


def test_logistic_closed_form_at_zero():
    X, y = _rand_logit(500, 7, 1)
    d = O.GlmData(O.LOGIT, X, y)
    for mode in (0, 1):
        st, lp, g = d.lpgrad(np.zeros(8), mode=mode)
        assert st == 0
        assert lp == pytest.approx(-500 * math.log(2.0), rel=1e-14)
        np.testing.assert_allclose(g[1:], X.T @ (y - 0.5), rtol=1e-12, atol=1e-13)
        assert g[0] == pytest.approx(np.sum(y - 0.5), abs=1e-12)


def test_logistic_label_flip_invariance():
    X, y = _rand_logit(300, 5, 2)
    th = np.random.default_rng(3).standard_normal(6)
    a = O.GlmData(O.LOGIT, X, y)
    b = O.GlmData(O.LOGIT, X, 1 - y)
    _, lpa, ga = a.lpgrad(th)
    _, lpb, gb = b.lpgrad(-th)
    assert lpa == pytest.approx(lpb, rel=1e-14)
    np.testing.assert_allclose(ga, -gb, rtol=1e-13, atol=1e-13)


def test_normal_closed_form_at_truth():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((200, 4))
    beta = rng.standard_normal(4)
    alpha = 0.3
    y = X @ beta + alpha
    d = O.GlmData(O.NORMAL, X, y, prior=(0.0, 0.0, 0.0, 1.0))  # flat priors
    u = 0.7
    th = np.concatenate([[alpha], beta, [u]])
    st, lp, g = d.lpgrad(th, mode=1, jacobian=False)
    assert st == 0
    # residuals are O(1e-16) after rounding of y; SS ~ 0
    assert lp == pytest.approx(-200 * u, abs=1e-20 + 1e-12)
    np.testing.assert_allclose(g[:5], 0.0, atol=1e-12)
    assert g[5] == pytest.approx(-200.0, abs=1e-10)


def _central_fd(fun, th, h=1e-6):
    g = np.zeros_like(th)
    for i in range(th.size):
        e = np.zeros_like(th)
        e[i] = h
        g[i] = (fun(th + e) - fun(th - e)) / (2 * h)
    return g


@pytest.mark.parametrize("family", [O.LOGIT, O.NORMAL, O.HIER])
def test_gradient_matches_finite_differences(family):
    rng = np.random.default_rng(10 + family)
    if family == O.HIER:
        offsets = np.array([0, 5, 6, 14, 20])
        N, K = 20, 3
    else:
        offsets, N, K = None, 60, 4
    X = rng.standard_normal((N, K)) / 2
    y = rng.standard_normal(N) if family == O.NORMAL else (rng.random(N) < 0.5).astype(np.int32)
    d = O.GlmData(family, X, y, offsets=offsets)
    th = rng.standard_normal(d.D) * 0.5
    st, lp, g = d.lpgrad(th)
    assert st == 0
    fd = _central_fd(lambda t: d.lpgrad(t)[1], th)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6)


def _tape_lp(family, X, y, th, prior, offsets=None):
    N, K = X.shape
    v = [T.Var(t) for t in th]
    alpha, beta = v[0], v[1:1 + K]
    lp = T.Var(0.0)

    def nprior(x, s):
        return -0.5 * (x / s) * (x / s)

    if family == O.LOGIT:
        for n in range(N):
            eta = alpha
            for k in range(K):
                eta = eta + X[n, k] * beta[k]
            lp = lp + y[n] * eta - T.log1p(T.exp(eta))
    elif family == O.NORMAL:
        u = v[K + 1]
        sigma = T.exp(u)
        for n in range(N):
            mu = alpha
            for k in range(K):
                mu = mu + X[n, k] * beta[k]
            r = (y[n] - mu) / sigma
            lp = lp - 0.5 * r * r - u
        lp = lp + nprior(sigma, prior[2]) + u
    else:
        tau = [T.exp(x) for x in v[1 + K:2 + 2 * K]]
        z = v[2 + 2 * K:]
        for j in range(len(offsets) - 1):
            zj = z[j * (K + 1):(j + 1) * (K + 1)]
            for n in range(offsets[j], offsets[j + 1]):
                eta = alpha + tau[0] * zj[0]
                for k in range(K):
                    eta = eta + X[n, k] * (beta[k] + tau[k + 1] * zj[k + 1])
                lp = lp + y[n] * eta - T.log1p(T.exp(eta))
        for k in range(K + 1):
            lp = lp + nprior(tau[k], prior[3]) + v[1 + K + k]
        for zi in z:
            lp = lp - 0.5 * zi * zi
    if prior[0] > 0:
        lp = lp + nprior(alpha, prior[0])
    for b in beta:
        lp = lp + nprior(b, prior[1])
    return lp.val, T.grad(lp, v)


@pytest.mark.parametrize("family", [O.LOGIT, O.NORMAL, O.HIER])
def test_gradient_matches_reverse_mode_tape(family):
    rng = np.random.default_rng(20 + family)
    offsets = np.array([0, 3, 4, 9]) if family == O.HIER else None
    N, K = (9, 2) if family == O.HIER else (7, 3)
    X = rng.standard_normal((N, K))
    y = rng.standard_normal(N) if family == O.NORMAL else (rng.random(N) < 0.5).astype(np.int32)
    prior = (5.0, 2.0, 2.0, 1.0)
    d = O.GlmData(family, X, y, offsets=offsets, prior=prior)
    th = rng.standard_normal(d.D) * 0.7
    st, lp, g = d.lpgrad(th, mode=1)
    lp_t, g_t = _tape_lp(family, X, y, th, prior, offsets)
    assert st == 0
    assert lp == pytest.approx(lp_t, rel=1e-13)
    np.testing.assert_allclose(g, g_t, rtol=1e-12, atol=1e-13)


def test_propto_false_adds_constants():
    X, y = _rand_logit(100, 3, 5)
    d = O.GlmData(O.LOGIT, X, y, prior=(5.0, 2.0, 2.0, 1.0))
    th = np.array([0.1, -0.2, 0.3, 0.05])
    _, lp1, _ = d.lpgrad(th, propto=True)
    _, lp0, _ = d.lpgrad(th, propto=False)
    const = -(math.log(5.0) + 0.5 * math.log(2 * math.pi)) - 3 * (
        math.log(2.0) + 0.5 * math.log(2 * math.pi))
    assert lp0 - lp1 == pytest.approx(const, rel=1e-13)


def test_two_pass_and_compensated_agree():
    X, y = _rand_logit(20000, 30, 6)
    d = O.GlmData(O.LOGIT, X, y)
    th = np.random.default_rng(7).standard_normal(31)
    _, lp0, g0 = d.lpgrad(th, mode=0)
    _, lp1, g1 = d.lpgrad(th, mode=1)
    assert abs(lp0 - lp1) / abs(lp1) < 1e-13
    assert np.linalg.norm(g0 - g1) / np.linalg.norm(g1) < 1e-12


def test_domain_error_on_nonfinite_linear_predictor():
    X, y = _rand_logit(10, 2, 8)
    d = O.GlmData(O.LOGIT, X, y)
    st, _, _ = d.lpgrad(np.array([np.inf, 0.0, 0.0]))
    assert st == 1


# ---------------- oracle sampler ----------------

def test_oracle_sampler_gauss_moments():
    f = O.model_fn("orc_gauss_model_cb")
    draws = []
    for chain in range(1, 5):
        rc, d, eps, minv, ng = O.hmc_nuts_diag_e_adapt(f, None, 3, seed=1234, chain=chain,
                                                       num_warmup=500, num_samples=1000)
        assert rc == 0
        assert d.shape == (1000, 10)
        draws.append(d[:, 7:])
    x = np.stack(draws, axis=1)  # iterations x chains x params
    for p in range(3):
        m = x[:, :, p].mean()
        assert abs(m) < 4 * analyze.mcse_mean(x[:, :, p])
        v = x[:, :, p].var()
        assert abs(v - 1.0) < 0.15


def test_oracle_sampler_output_shape_and_thinning():
    f = O.model_fn("orc_gauss_model_cb")
    rc, d, _, _, _ = O.hmc_nuts_diag_e_adapt(f, None, 2, num_warmup=200, num_samples=400,
                                             num_thin=5, save_warmup=1)
    assert rc == 0
    assert d.shape == ((200 + 400) // 5, 9)


def test_cpu_baseline_leg_agrees_with_oracle():
    """The benchmark's timed CPU leg (oracle/glm_cpu_baseline.c, -O3 OpenMP)
    computes the same density as the oracle; threads only regroup the sums."""
    import ctypes as Cc
    import os
    lib = Cc.CDLL(os.path.join(os.path.dirname(O.__file__), "libcpu_baseline.so"))
    f = lib.cpu_logit_lpgrad
    f.argtypes = [Cc.c_void_p, Cc.c_void_p, Cc.c_int64, Cc.c_int64, Cc.c_double, Cc.c_double,
                  Cc.c_void_p, Cc.c_int, Cc.POINTER(Cc.c_double), Cc.c_void_p]
    X, y = _rand_logit(3001, 23, 5)
    rng = np.random.default_rng(6)
    th = rng.standard_normal(24)
    ref = O.GlmData(O.LOGIT, X, y, prior=(5.0, 2.0, 2.0, 1.0))
    st, lp_r, g_r = ref.lpgrad(th, mode=1)
    for threads in (1, 3):
        g = np.zeros(24)
        lp = Cc.c_double()
        assert f(X.ctypes.data, y.ctypes.data, 3001, 23, 5.0, 2.0, th.ctypes.data, threads,
                 Cc.byref(lp), g.ctypes.data) == 0
        assert abs(lp.value - lp_r) <= 1e-12 * abs(lp_r)
        assert np.linalg.norm(g - g_r) <= 1e-12 * np.linalg.norm(g_r)
