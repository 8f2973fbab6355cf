"""Host-side C++ checks (no GPU): the lean FP64 link functions the GLM
kernels use, compiled for the host from the same header, against the C
library."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def sx_math_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cpp") / "sx_math_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "cpp", "sx_math_check.cpp")], check=True)
    return exe


def test_lean_link_math_accuracy(sx_math_check):
    out = subprocess.run([sx_math_check], check=True, capture_output=True, text=True).stdout
    vals = dict(zip(out.split()[0::2], map(float, out.split()[1::2])))
    # exp / log1p / reciprocal within a few ulps of the C library; the row
    # term and residual of the logistic link within 1e-15 relative
    assert vals["exp_ulps"] <= 4.0, out
    assert vals["log1p_ulps"] <= 4.0, out
    assert vals["rcp_ulps"] <= 1.0, out
    assert vals["term_rel"] <= 1e-15, out
    assert vals["r_rel"] <= 1e-15, out


def test_staged_tile_layout_matches_mfma_operands(tmp_path):
    """The LDS tile layout of the BF16 MFMA passes (staging, element
    writes, chunk copies from the hierarchical pass's split copy of X,
    forward reads and transposing adjoint reads of
    stan_amd/csrc/tile_layout.hpp), compiled for the host, delivers every
    operand element the MFMA expects, for every padded K up to 256 and
    every hierarchical K up to 31."""
    exe = str(tmp_path / "tile_layout_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe,
                    os.path.join(HERE, "cpp", "tile_layout_check.cpp")], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
