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


def _csv_rows(path):
    """names and numeric rows of a CSV with '#' comment lines"""
    names, rows = None, []
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        if names is None:
            names = line.strip().split(",")
        else:
            rows.append([float(v) for v in line.strip().split(",")])
    return names, np.array(rows)


def test_diagnostic_file_and_metric_json(tmp_path):
    """The reference's diagnostic_writer and metric_writer outputs
    (mcmc_writer.hpp:157-190; base_hmc.hpp:70-76): per saved draw the
    unconstrained position, momentum and potential gradient of the selected
    point, and one JSON record of the adapted tuning parameters.  For the
    standard normal model the potential gradient equals the position, and
    the selected point is the draw."""
    import json
    f = O.model_fn("orc_gauss_model_cb")
    kw = dict(random_seed=33, chain=1, num_warmup=200, num_samples=150, save_warmup=1)
    paths = dict(sample_csv=tmp_path / "s.csv", diagnostic_csv=tmp_path / "d.csv",
                 metric_json=tmp_path / "m.json")
    out = S.hmc_nuts_diag_e_adapt_cb_files(f, None, 3, sig_figs=18, **paths, **kw)
    assert out["rc"] == 0
    table = S.hmc_nuts_diag_e_adapt_cb(f, None, 3, **kw)
    np.testing.assert_array_equal(out["draws"], table["draws"])  # outputs do not touch the RNG
    names, rows = _csv_rows(paths["diagnostic_csv"])
    assert names == ["lp__", "accept_stat__", "stepsize__", "treedepth__", "n_leapfrog__",
                     "divergent__", "energy__", "q.1", "q.2", "q.3", "p_q.1", "p_q.2", "p_q.3",
                     "g_q.1", "g_q.2", "g_q.3"]
    assert rows.shape == (350, 16)
    np.testing.assert_array_equal(rows[:, :7], table["draws"][:, :7])
    q, p, g = rows[:, 7:10], rows[:, 10:13], rows[:, 13:16]
    np.testing.assert_array_equal(q, table["draws"][:, 7:])
    np.testing.assert_array_equal(g, q)  # V = |q|^2 / 2
    # energy__ = V + T at the selected point: T = p . (M^-1 p) / 2; the
    # sampling-phase metric is the adapted one
    minv = table["inv_metric"]
    e = 0.5 * (q ** 2).sum(1) + 0.5 * (p * p * minv).sum(1)
    np.testing.assert_allclose(rows[200:, 6], e[200:], rtol=1e-12)
    rec = json.loads(paths["metric_json"].read_text())
    assert rec["metric_type"] == "diag_e"
    assert rec["stepsize"] == table["stepsize"]
    assert rec["inv_metric"] == list(table["inv_metric"])
    # the sample CSV is the one the _csv entry point writes
    r = read_stan_csv(paths["sample_csv"])
    np.testing.assert_array_equal(r["samples"], table["draws"])
