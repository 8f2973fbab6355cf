"""Contexts created from Stan JSON data files equal contexts created from
the same arrays (bitwise), and the hierarchical model's rows are grouped
by the file's group index."""
import json
import os

import numpy as np
import pytest

import oracle as O
import stan_amd as S
from stan_amd import workloads as W

pytestmark = pytest.mark.gpu

GLM_JSON = os.path.join(os.path.dirname(__file__), "golden", "normal_glm_test.json")
BRMS_PRIOR = (3.0, 3.0, 2.0, 0.0, 1.0)


def test_normal_glm_from_reference_json():
    d = json.load(open(GLM_JSON))
    X, Y = np.array(d["X"]), np.array(d["Y"])
    mj = S.glm_from_json(GLM_JSON, S.NORMAL, prior=BRMS_PRIOR)
    ma = S.GlmModel(S.NORMAL, X, Y, prior=BRMS_PRIOR)
    ref = O.GlmData(O.NORMAL, X, Y, prior=BRMS_PRIOR)
    assert mj.D == 7
    for th in W.random_thetas(7, 4, seed=5) * 0.5:
        lpj, gj = mj.log_prob_grad(th)
        lpa, ga = ma.log_prob_grad(th)
        assert lpj == lpa
        np.testing.assert_array_equal(gj, ga)
        st, lpr, gr = ref.lpgrad(th, mode=1)
        assert abs(lpj - lpr) <= 1e-12 * abs(lpr)
        assert np.linalg.norm(gj - gr) <= 1e-12 * np.linalg.norm(gr)
    mj.close()
    ma.close()


def test_hier_from_json_groups_rows(tmp_path):
    d = W.hier_logistic(J=7, K=3, seed=55, mean_log_size=np.log(40))
    g = np.repeat(np.arange(1, 8), np.diff(d["offsets"]))
    perm = np.random.Generator(np.random.PCG64(1)).permutation(len(g))
    p = tmp_path / "hier.json"
    p.write_text(json.dumps(dict(N=len(g), K=3, J=7, X=d["X"][perm].tolist(),
                                 y=d["y"][perm].tolist(), ids=g[perm].tolist())))
    mj = S.glm_from_json(p, S.HIER_LOGIT, prior=d["prior"])
    ref = O.GlmData(O.HIER, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"])
    for th in W.random_thetas(mj.D, 3, seed=6) * 0.5:
        lp, gr = mj.log_prob_grad(th)
        st, lpr, grr = ref.lpgrad(th, mode=1)
        assert abs(lp - lpr) <= 1e-12 * abs(lpr)
        assert np.linalg.norm(gr - grr) <= 1e-12 * np.linalg.norm(grr)
    mj.close()
