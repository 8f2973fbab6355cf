"""The model boundary in the generated style, on the CPU: a model deriving
from model_base_crtp<M> with one log_prob template over double and the
library's reverse-mode var (stanx/var.hpp, stanx/model.hpp), compiled
against the product headers (tests/cpp/model_crtp_check.cpp).

Pins: the tape gradient against the closed form of the same density; the
eight log_prob overloads (model_base.hpp:191-302, 404-544) against each
other and the closed-form Jacobian and normalising constants; a
precomputed-gradient entry inside a larger expression against its closed
form; the tape is cleared after the model throws; the adaptive service
over the model recovers the posterior of mu (mean ~ ybar, sd ~ sigma /
sqrt(N))."""
import math
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def crtp(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cpp") / "model_crtp_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "cpp", "model_crtp_check.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return dict(zip(out.split()[0::2], map(float, out.split()[1::2])))


def test_tape_gradient_matches_closed_form(crtp):
    assert crtp["crtp_lp_rel"] <= 1e-14
    assert crtp["crtp_grad_rel"] <= 1e-13


def test_overloads_and_constants(crtp):
    # jacobian adds u = 0.4 (sigma = exp(u)); the normalising constants of
    # 100 normal rows and the two priors (half-normal on sigma: + log 2)
    assert crtp["jacobian_term"] == pytest.approx(0.4, abs=1e-12)
    assert crtp["propto_jacobian_term"] == pytest.approx(0.4, abs=1e-12)
    c = -math.log(10) - math.log(5) + math.log(2) - 102 * 0.5 * math.log(2 * math.pi)
    assert crtp["constant_term"] == pytest.approx(c, rel=1e-13)
    assert crtp["var_vs_double_maxdiff"] == 0.0


def test_precomputed_gradient_entry(crtp):
    assert crtp["precomputed_node_err"] == 0.0


def test_tape_cleared_after_throw(crtp):
    assert crtp["threw"] == 1 and crtp["tape_after_throw"] == 0


def test_service_over_generated_style_model(crtp):
    assert crtp["service_rc"] == 0 and crtp["service_rows"] == 4000
    assert abs(crtp["mu_mean"] - crtp["ybar"]) < 0.02
    assert crtp["mu_sd"] == pytest.approx(math.sqrt(2.0) / 10.0, rel=0.1)
