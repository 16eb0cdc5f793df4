"""The package's vectorised complementarity metric against the oracle's (same definition,
different summation order), with frictionless contacts, restitution and static sides."""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import metrics, scenes


@pytest.mark.parametrize("seed", range(3))
def test_complementarity_metric_matches_oracle(seed):
    prob = scenes.random_problem(nb=60, nc=180, seed=seed, warm=True, p_point=0.0)
    rng = np.random.default_rng(seed)
    prob.rest[:] = np.where(rng.random(prob.rest.shape) < 0.5, rng.uniform(0, 0.8, prob.rest.shape), 0.0)
    prob.mu[rng.random(prob.nc) < 0.2] = 0.0
    assert (prob.body_a < 0).any() and (prob.mu == 0).any() and (prob.rest > 0).any()
    for k in (1, 25):
        st, v, imp, res = oracle.pj_solve(prob, 1e-3, oracle.params(max_iter=k, tol_abs=0, tol_rel=0))
        a = oracle.complementarity(prob, v, imp)
        assert a > 0
        assert metrics.complementarity(prob, v, imp) == pytest.approx(a, rel=1e-12)


def test_complementarity_vanishes_at_convergence():
    """Bodies each held by static contacts: the residual falls to round-off as PJ converges, and
    is far from zero after one sweep."""
    prob = scenes.random_problem(nb=6, nc=6, seed=4, p_point=0.0, p_static=1.0, warm=False)
    one = oracle.pj_solve(prob, 1e-3, oracle.params(max_iter=1, tol_abs=0, tol_rel=0))
    conv = oracle.pj_solve(prob, 1e-3, oracle.params(max_iter=20000, tol_abs=1e-15, tol_rel=1e-15))
    assert np.abs(conv[2][:, 0]).max() > 0
    assert metrics.complementarity(prob, one[1], one[2]) > 1e-3
    assert metrics.complementarity(prob, conv[1], conv[2]) < 1e-9
