"""Stan CSV files: the sample output of the sampler services.

read_stan_csv restates the reader of src/stan/io/stan_csv_reader.hpp
(structs at :32-95, read_metadata :104-196, read_header :198-231,
read_adaptation :233-284, read_samples :286-348): the configuration
comments before the header become metadata, the header gives the column
names, the comment block after the header gives the adapted step size and
inverse metric (one line of diagonal elements for diag_e, one line per
row for dense_e), numeric lines are the draws, and the trailing comment
block gives the elapsed times.  It reads files written by CmdStan and by
sx_hmc_nuts_diag_e_adapt_csv alike.
"""
import re

import numpy as np

_META_KEYS = {
    "stan_version_major": int, "stan_version_minor": int, "stan_version_patch": int,
    "model": str, "num_samples": int, "num_warmup": int, "save_warmup": None, "thin": int,
    "id": int, "seed": int, "max_depth": int, "algorithm": str, "engine": str, "method": str,
    "init": str, "metric": str, "stepsize": float, "delta": float,
}


def _value(kind, text):
    text = text.replace("(Default)", "").strip()
    if kind is None:  # boolean, as "true"/"false" or 1/0
        return text in ("true", "1")
    if kind is str:
        return text
    try:
        return kind(text)
    except ValueError:
        return text


def read_stan_csv(path):
    """Returns dict(metadata, column_names, samples (rows x columns),
    step_size, inv_metric, timing={warmup, sampling, total})."""
    with open(path) as f:
        lines = f.read().splitlines()
    meta, names, rows = {}, None, []
    step_size, metric, timing = None, None, {}
    expect_metric = 0  # metric lines still to read
    dense_rows = []
    for line in lines:
        if line.startswith("#"):
            body = line[1:].strip()
            if names is None:
                m = re.match(r"^([A-Za-z_][A-Za-z_0-9]*)\s*=\s*(.*)$", body)
                if m and m.group(1) in _META_KEYS:
                    meta[m.group(1)] = _value(_META_KEYS[m.group(1)], m.group(2))
                continue
            if expect_metric and body:
                row = np.array([float(v) for v in body.split(",")])
                if expect_metric == -1:  # diagonal: one line
                    metric = row
                    expect_metric = 0
                else:  # dense: as many lines as the first row has entries
                    dense_rows.append(row)
                    if len(dense_rows) == len(dense_rows[0]):
                        metric = np.stack(dense_rows)
                        expect_metric = 0
                continue
            m = re.match(r"^Step size = (.*)$", body)
            if m:
                step_size = float(m.group(1))
                continue
            if body.startswith("Diagonal elements of inverse mass matrix"):
                expect_metric = -1
                continue
            if body.startswith("Elements of inverse mass matrix"):
                expect_metric = 1
                dense_rows = []
                continue
            m = re.match(r"^(?:Elapsed Time:)?\s*([0-9.eE+-]+) seconds \((Warm-up|Sampling|Total)\)",
                         body)
            if m:
                timing[{"Warm-up": "warmup", "Sampling": "sampling",
                        "Total": "total"}[m.group(2)]] = float(m.group(1))
            continue
        if not line.strip():
            continue
        if names is None:
            names = line.split(",")
            continue
        rows.append([float(v) for v in line.split(",")])
    samples = np.array(rows) if rows else np.zeros((0, len(names or [])))
    return dict(metadata=meta, column_names=names, samples=samples, step_size=step_size,
                inv_metric=metric, timing=timing)
