"""Warm-up adaptation of the product's C++ sampler (stanx/mcmc/adaptation.hpp,
stanx/mcmc/warmup_plan.hpp), compiled for the host:

- dual averaging against the reference's goldens
  (src/test/unit/mcmc/stepsize_adaptation_test.cpp:59-143), on the product
  (the oracle's copy is checked in test_oracle.py);
- the regularised zero-variance window
  (src/test/unit/mcmc/var_adaptation_test.cpp:5-27) on product and oracle;
- the window schedule against the schedule the reference's documentation
  states for 1000 warm-up iterations (windows 75 / 25, 50, 100, 200, 500 / 50)
  and against a statement of the reference's window counter
  (src/stan/mcmc/windowed_adaptation.hpp:23-110) over many configurations,
  including its two corner cases (num_warmup < 20; buffers that do not fit).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
U32 = 0xFFFFFFFF


@pytest.fixture(scope="module")
def check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cpp") / "adaptation_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "cpp", "adaptation_check.cpp")], check=True)

    def run(lines):
        out = subprocess.run([exe], input="\n".join(lines) + "\n", check=True,
                             capture_output=True, text=True).stdout
        return [ln.split() for ln in out.strip().splitlines()]
    return run


def reference_window_counter(nw, ib, tb, bw, calls):
    """The reference's windowed_adaptation in 32-bit unsigned arithmetic:
    the constructor restarts with zeros; set_window_params either returns
    early (nw < 20), sets the 15/75/10 split without restarting, or sets the
    given buffers and restarts.  Returns (fed iterations, closing iterations)
    over `calls` calls of learn_variance."""
    n_w, i_b, t_b, b_w = 0, 0, 0, 0
    size = b_w
    nxt = (i_b + size - 1) & U32
    if nw >= 20:
        if ib + bw + tb > nw:
            n_w = nw
            i_b = int(0.15 * nw)
            t_b = int(0.10 * nw)
            b_w = nw - (i_b + t_b)
        else:
            n_w, i_b, t_b, b_w = nw, ib, tb, bw
            size = b_w
            nxt = (i_b + size - 1) & U32
    last = (n_w - t_b - 1) & U32
    fed, closes = [], []
    for counter in range(calls):
        if counter >= i_b and counter < ((n_w - t_b) & U32) and counter != n_w:
            fed.append(counter)
        if counter == nxt and counter != n_w:
            closes.append(counter)
            if nxt != last:
                size = (size * 2) & U32
                nxt = (counter + size) & U32
                if nxt != last:
                    boundary = (nxt + 2 * size) & U32
                    if boundary >= ((n_w - t_b) & U32):
                        nxt = last
    return fed, closes


def _plan(check, cfgs):
    out = check([f"plan {nw} {ib} {tb} {bw}" for nw, ib, tb, bw in cfgs])
    res = []
    for r in out:
        fb, fe, n = int(r[1]), int(r[2]), int(r[3])
        res.append((fb, fe, [int(x) for x in r[4:4 + n]]))
    return res


def test_default_schedule_of_1000_warmup(check):
    (fb, fe, closes), = _plan(check, [(1000, 75, 50, 25)])
    # windows [75, 99], [100, 149], [150, 249], [250, 449], [450, 949]
    assert (fb, fe) == (75, 950)
    assert closes == [99, 149, 249, 449, 949]


def test_schedule_matches_reference_counter(check):
    rng = np.random.default_rng(5)
    cfgs = [(10, 75, 50, 25), (19, 1, 1, 1), (20, 3, 2, 5), (100, 75, 50, 25),
            (149, 75, 50, 25), (150, 75, 50, 25), (151, 75, 50, 25), (2000, 75, 50, 25),
            (300, 0, 0, 10), (50, 0, 0, 10), (500, 10, 0, 1), (400, 7, 3, 0), (1000, 0, 0, 1)]
    for _ in range(200):
        nw = int(rng.integers(0, 3000))
        cfgs.append((nw, int(rng.integers(0, 200)), int(rng.integers(0, 200)),
                     int(rng.integers(0, 200))))
    for cfg, (fb, fe, closes) in zip(cfgs, _plan(check, cfgs)):
        nw = cfg[0]
        fed, ref_closes = reference_window_counter(*cfg, calls=nw)
        assert closes == ref_closes, cfg
        assert [i for i in range(nw) if fb <= i < fe] == fed, cfg


def test_reduced_split_feeds_but_never_closes(check):
    # 100 < 75 + 25 + 50: 15 / 75 / 10, fed over [15, 90), no window closes
    (fb, fe, closes), = _plan(check, [(100, 75, 50, 25)])
    assert (fb, fe, closes) == (15, 90, [])


def test_dual_averaging_goldens_on_product(check, goldens):
    g = goldens["learn_stepsize"]
    fp = g["fixed_point"]
    lines = [f"tune {fp['mu']} {fp['delta']} 0.05 0.75 10 0 0 0 {fp['stat']}"]
    for case in g["cases"]:
        n, sb, xb, mu, delta, gamma, kappa, t0 = case["state"]
        lines.append(f"tune {mu!r} {delta!r} {gamma!r} {kappa!r} {t0!r} {n!r} {sb!r} {xb!r} "
                     f"{case['stat']!r}")
    out = check(lines)
    assert float(out[0][1]) == fp["eps_out"]
    for case, r in zip(g["cases"], out[1:]):
        eps, n, sb, xb = map(float, r[1:])
        assert abs(eps - case["eps_out"]) < case["tol"]
        if "counter" in case:
            assert n == case["counter"]
            assert abs(sb - case["s_bar"]) < 1e-14
            assert abs(xb - case["x_bar"]) < 1e-14


def test_zero_variance_window_golden(check, goldens):
    g = goldens["var_adaptation_zero"]
    out = check([f"varzero {g['n']} {g['n_learn']} {g['num_warmup']} 0 0 10"])
    vals = [float(v) for v in out[0][1:]]
    assert vals == [g["expected"]] * g["n"]  # EXPECT_EQ in the reference test


def test_zero_variance_window_golden_on_oracle(goldens):
    g = goldens["var_adaptation_zero"]
    n = g["n"]
    f = O.lib().orc_var_window_zero
    f.argtypes = [C.c_int, C.c_int, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_void_p]
    f.restype = C.c_int
    out = np.full(n, -1.0)
    assert f(n, g["n_learn"], g["num_warmup"], 0, 0, 10, out.ctypes.data) == 0
    assert list(out) == [g["expected"]] * n
