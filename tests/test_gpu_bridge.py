"""The GPU log density as one precomputed-gradient entry inside a model
written in the generated style (tests/cpp/gpu_bridge_check.cpp, built by
__graft_entry__.build() against stan_amd/libstanx.so): the composite
density's value and gradient against sx_glm_logp_grad plus the closed form
of the extra term (1e-12), and the adaptive service over it."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "gpu_bridge_check")


def test_gpu_glm_inside_generated_style_model():
    print("gpu bridge driver", flush=True)
    assert os.path.exists(EXE), "run __graft_entry__.build() first"
    out = subprocess.run([EXE], check=True, capture_output=True, text=True, timeout=120).stdout
    v = dict(zip(out.split()[0::2], map(float, out.split()[1::2])))
    assert v["bridge_lp_rel"] <= 1e-12, out
    assert v["bridge_grad_rel"] <= 1e-12, out
    assert v["service_rc"] == 0 and v["service_rows"] == 200, out
