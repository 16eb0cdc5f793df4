"""GPU parity of the Anderson-mixed projected Jacobi loop (anderson_loop, projected_jacobi.cpp:
224-310) against the oracle.

The device forms the least-squares coefficients from the Gram matrix of the residual differences
(dot products reduced on the device, a pivoted Cholesky with the pivot rule of Eigen's
ColPivHouseholderQR) where the oracle runs the column-pivoted Householder QR itself, so the two
agree to rounding, not bitwise -- and Anderson mixing amplifies rounding: a relative change of
1e-15 in the initial velocities moves the oracle's own result by up to ~1e-5 after 64 sweeps on
these problems, and changes the iteration count to tolerance on most of them.  The bar is
therefore 1e-10 or 100 times the oracle's own response to such a perturbation of the same run,
whichever is larger (1e-10 itself up to 9 sweeps); runs to tolerance use problems whose count is
unchanged by the perturbation; on affine problems the exact fixed point after the sweeps the
Krylov argument fixes (see test_oracle.py) is checked to 1e-12."""
import copy

import numpy as np
import pytest

import oracle
from cardillocxx_amd import native, scenes
from test_oracle import G, plane_scene, two_planes

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    a = np.asarray(a, float).ravel()
    b = np.asarray(b, float).ravel()
    return np.abs(a - b).max(initial=0.0) / max(np.abs(b).max(initial=0.0), 1e-300)


def device(ctx, prob, dt=1e-3, **kw):
    p = native.default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    p.accel = native.CDX_ACCEL_ANDERSON
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    info = ctx.pj_solve(dt, 1.0, p)
    return ctx.velocities(), ctx.impulses(), info


def oracle_run(prob, dt=1e-3, **kw):
    kw = dict(kw)
    kw.setdefault("anderson_m", 2)
    st, v, imp, res = oracle.pj_solve(prob, dt, oracle.params(anderson=1, **kw))
    assert st == 0
    return v, imp, res


def bar(prob, v_ref, imp_ref, **kw):
    """max(TOL, 100 x the oracle's response to a relative change of +-1e-15 in the velocities)"""
    s = 0.0
    for sign in (1, -1):
        pert = copy.deepcopy(prob)
        pert.v = pert.v * (1 + sign * 1e-15)
        v2, imp2, _ = oracle_run(pert, **kw)
        s = max(s, rel(v2, v_ref), rel(imp2, imp_ref))
    return max(TOL, 100 * s)


CASES = [
    dict(nb=20, nc=60, seed=2),
    dict(nb=50, nc=200, seed=3, p_point=0.3),
    dict(nb=200, nc=300, seed=4, rest=0.4),
    dict(nb=300, nc=450, seed=5, p_frictionless=0.0, p_static=0.3),
    dict(nb=3, nc=8, seed=6, p_static=0.3, p_point=1.0, warm=False),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"nb{c['nb']}_nc{c['nc']}_s{c['seed']}")
@pytest.mark.parametrize("sweeps", [1, 2, 5, 9, 64])
@pytest.mark.parametrize("m", [1, 2, 5])
def test_anderson_fixed_sweeps(ctx, case, sweeps, m):
    prob = scenes.random_problem(**case)
    kw = dict(max_iter=sweeps, tol_abs=0.0, tol_rel=0.0, anderson_m=m)
    v_ref, imp_ref, res = oracle_run(prob, **kw)
    v, imp, info = device(ctx, prob, **kw)
    assert info.iters == res.iters == sweeps
    b = bar(prob, v_ref, imp_ref, **kw)
    if sweeps <= 9:
        assert b == TOL  # short runs: the plain rounding bar
    assert rel(v, v_ref) <= b and rel(imp, imp_ref) <= b


# iteration counts to tolerance are unchanged by the perturbation on these two (not on the
# others, where an ulp in the velocities changes the count)
STABLE = [CASES[2], dict(nb=60, nc=100, seed=13, p_frictionless=1.0)]


@pytest.mark.parametrize("case", STABLE, ids=lambda c: f"nb{c['nb']}_nc{c['nc']}_s{c['seed']}")
@pytest.mark.parametrize("tol", [1e-5, 1e-9])
def test_anderson_to_tolerance(ctx, case, tol):
    prob = scenes.random_problem(**case)
    kw = dict(tol_abs=tol, tol_rel=tol)
    v_ref, imp_ref, res = oracle_run(prob, **kw)
    v, imp, info = device(ctx, prob, **kw)
    assert info.iters == res.iters and res.residual <= 1 and info.residual <= 1
    b = bar(prob, v_ref, imp_ref, **kw)
    assert rel(v, v_ref) <= b and rel(imp, imp_ref) <= b
    std = oracle.pj_solve(prob, 1e-3, oracle.params(**kw))
    assert info.iters < std[3].iters  # fewer sweeps than the standard loop on these


def test_anderson_exact_on_affine_sweeps(ctx):
    """Two coupled frictionless contacts: m = 2 gives the exact impulses after 3 sweeps; one
    contact: m = 1 after 2 (the Krylov argument of the oracle test)."""
    prob, W, pstar = two_planes()
    for m in (2, 3):
        v, imp, info = device(ctx, prob, max_iter=3, tol_abs=0.0, tol_rel=0.0, warmstart=0, anderson_m=m)
        assert info.iters == 3
        np.testing.assert_allclose(imp[:, 0], pstar, rtol=0, atol=1e-12)
        assert np.abs(W @ v).max() < 1e-12
    one = plane_scene().problem()
    one.v[2] = -1.0
    v, imp, info = device(ctx, one, max_iter=2, tol_abs=0.0, tol_rel=0.0, warmstart=0, anderson_m=1)
    assert imp[0, 0] == pytest.approx(1.0 + G * 1e-3, rel=1e-13)
    # run to tolerance: converged on the sweep after the exact one
    v, imp, info = device(ctx, prob, tol_abs=1e-12, tol_rel=1e-12, warmstart=0, anderson_m=2)
    assert info.iters == 4 and info.residual <= 1


@pytest.mark.parametrize("sweeps", [1, 7, 30])
def test_anderson_with_block_preconditioner(ctx, sweeps):
    prob = scenes.random_problem(nb=40, nc=90, seed=8, p_point=0.2)
    kw = dict(max_iter=sweeps, tol_abs=0.0, tol_rel=0.0, block_precond=1)
    v_ref, imp_ref, res = oracle_run(prob, **kw)
    v, imp, info = device(ctx, prob, **kw)
    assert info.iters == res.iters == sweeps
    b = bar(prob, v_ref, imp_ref, **kw)
    assert rel(v, v_ref) <= b and rel(imp, imp_ref) <= b


def test_anderson_graph_replay_and_refusals(ctx):
    """Graph-replayed chunks (8 iterations each) give the direct launches' result (profiling
    launches every kernel directly) bitwise; the history window must be 1..8."""
    prob = scenes.random_problem(nb=50, nc=200, seed=3, p_point=0.3)
    kw = dict(max_iter=40, tol_abs=0.0, tol_rel=0.0)
    v1, imp1, _ = device(ctx, prob, **kw)
    ctx.set_profiling(True)
    v2, imp2, _ = device(ctx, prob, **kw)
    ctx.set_profiling(False)
    np.testing.assert_array_equal(v1, v2)
    np.testing.assert_array_equal(imp1, imp2)
    for m in (0, 9):
        with pytest.raises(native.CdxError):
            device(ctx, prob, max_iter=5, anderson_m=m)
