"""Stan CSV output and reader (the sample file of the services).

Pins: the reader on CmdStan's own output file (tests/golden/bern1.csv,
written by CmdStan 2.35.0, the fixture of the reference's analysis tests);
the writer by a round trip through the reader against the draws the same
service run returns as a table, and by the reference's line formats
(src/stan/services/util/mcmc_writer.hpp:91-238)."""
import numpy as np

import oracle as O
import stan_amd as S
from stan_amd.stan_csv import read_stan_csv

import os
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_reader_on_cmdstan_output():
    r = read_stan_csv(os.path.join(GOLDEN, "bern1.csv"))
    assert r["column_names"] == ["lp__", "accept_stat__", "stepsize__", "treedepth__",
                                 "n_leapfrog__", "divergent__", "energy__", "theta"]
    assert r["samples"].shape == (1000, 8)
    np.testing.assert_array_equal(r["samples"][0], [-6.80778, 0.991604, 0.883328, 2, 3, 0,
                                                    6.81551, 0.208532])
    assert r["step_size"] == 0.883328
    np.testing.assert_array_equal(r["inv_metric"], [0.512256])
    assert r["timing"] == {"warmup": 0.003, "sampling": 0.009, "total": 0.012}
    md = r["metadata"]
    assert md["seed"] == 86520 and md["num_samples"] == 1000 and md["max_depth"] == 10
    assert md["save_warmup"] is False and md["stan_version_minor"] == 35


def test_writer_round_trip(tmp_path):
    f = O.model_fn("orc_gauss_model_cb")
    path = tmp_path / "gauss.csv"
    kw = dict(random_seed=21, chain=2, num_warmup=150, num_samples=120, save_warmup=1,
              num_thin=2)
    out = S.hmc_nuts_diag_e_adapt_cb_csv(f, None, 3, str(path), sig_figs=18, **kw)
    assert out["rc"] == 0
    table = S.hmc_nuts_diag_e_adapt_cb(f, None, 3, **kw)
    np.testing.assert_array_equal(out["draws"], table["draws"])
    r = read_stan_csv(path)
    assert r["column_names"][:7] == ["lp__", "accept_stat__", "stepsize__", "treedepth__",
                                     "n_leapfrog__", "divergent__", "energy__"]
    assert len(r["column_names"]) == 10
    np.testing.assert_array_equal(r["samples"], table["draws"])  # 18 digits: exact
    assert r["samples"].shape == ((150 + 120) // 2, 10)
    # the adaptation lines are formatted by their own string streams, as in
    # the reference (six significant digits whatever sig_figs is)
    np.testing.assert_allclose(r["step_size"], table["stepsize"], rtol=5e-6)
    np.testing.assert_allclose(r["inv_metric"], table["inv_metric"], rtol=5e-6)
    assert set(r["timing"]) == {"warmup", "sampling", "total"}
    md = r["metadata"]
    assert md["seed"] == 21 and md["id"] == 2 and md["num_warmup"] == 150
    assert md["save_warmup"] is True and md["thin"] == 2
    text = path.read_text().splitlines()
    i = text.index("# Adaptation terminated")
    assert text[i + 1].startswith("# Step size = ")
    assert text[i + 2] == "# Diagonal elements of inverse mass matrix:"
    assert text[-5] == "# " and text[-1] == "# "
    assert text[-4].startswith("#  Elapsed Time: ") and text[-4].endswith(" seconds (Warm-up)")
    assert text[-3].startswith("#                ") and text[-3].endswith(" seconds (Sampling)")


def test_default_precision_is_six_digits(tmp_path):
    f = O.model_fn("orc_gauss_model_cb")
    path = tmp_path / "g6.csv"
    out = S.hmc_nuts_diag_e_adapt_cb_csv(f, None, 2, str(path), num_warmup=50, num_samples=20)
    r = read_stan_csv(path)
    np.testing.assert_allclose(r["samples"], out["draws"], rtol=5e-6, atol=1e-300)
