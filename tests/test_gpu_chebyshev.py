"""GPU parity of the Chebyshev-accelerated projected Jacobi loop (chebyshev_loop with
estimateSpectralRadius, projected_jacobi.cpp:155-228) against the oracle: velocities and
impulses to <= 1e-10 relative after the same sweeps, the same iteration count to tolerance,
through direct launches and graph-replayed chunks of sweeps."""
import copy

import numpy as np
import pytest

import oracle
from cardillocxx_amd import native, scenes
from test_oracle import orthogonal_point_mass

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    a = np.asarray(a, float).ravel()
    b = np.asarray(b, float).ravel()
    return np.abs(a - b).max(initial=0.0) / max(np.abs(b).max(initial=0.0), 1e-300)


def both(ctx, prob, dt=1e-3, **kw):
    st, v_ref, imp_ref, res = oracle.pj_solve(prob, dt, oracle.params(chebyshev=1, **kw))
    assert st == 0
    p = native.default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    p.accel = native.CDX_ACCEL_CHEBYSHEV
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    info = ctx.pj_solve(dt, 1.0, p)
    return (ctx.velocities(), ctx.impulses(), info), (v_ref, imp_ref, res)


CASES = [
    dict(nb=20, nc=60, seed=2),
    dict(nb=50, nc=200, seed=3, p_point=0.3),
    dict(nb=200, nc=300, seed=4, rest=0.4),
    dict(nb=300, nc=450, seed=5, p_frictionless=0.0, p_static=0.3),
    dict(nb=3, nc=8, seed=6, p_static=0.3, p_point=1.0, warm=False),  # full column rank: rho < 0.995
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"nb{c['nb']}_nc{c['nc']}_s{c['seed']}")
@pytest.mark.parametrize("sweeps", [1, 2, 9, 64])
def test_chebyshev_fixed_sweeps(ctx, case, sweeps):
    prob = scenes.random_problem(**case)
    (v, imp, info), (v_ref, imp_ref, res) = both(ctx, prob, max_iter=sweeps, tol_abs=0.0, tol_rel=0.0)
    assert info.iters == res.iters == sweeps
    assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL


# the loop converges on these (with rho clamped at 0.995 it does not on the others: the
# extrapolation overshoots the projection's active-set changes and the solve runs to max_iter)
CONVERGING = [CASES[0], CASES[4]]


@pytest.mark.parametrize("case", CONVERGING, ids=lambda c: f"nb{c['nb']}_nc{c['nc']}_s{c['seed']}")
@pytest.mark.parametrize("tol", [1e-5, 1e-9])
def test_chebyshev_to_tolerance_same_iterations(ctx, case, tol):
    prob = scenes.random_problem(**case)
    (v, imp, info), (v_ref, imp_ref, res) = both(ctx, prob, tol_abs=tol, tol_rel=tol)
    assert info.iters == res.iters and res.residual <= 1
    # the residual divides differences of size ~tol |v|: 1e-16 relative in v is ~1e-7 in it
    assert info.residual == pytest.approx(res.residual, rel=1e-5)
    assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL


@pytest.mark.parametrize("alpha", [0.5, 1.0])
def test_chebyshev_closed_form_rho(ctx, alpha):
    """W = I on one point mass: rho = 1 - alpha, and rho = 0 (power iteration stopped) at
    alpha = 1."""
    prob = orthogonal_point_mass()
    for kw in (dict(max_iter=5, tol_abs=0.0, tol_rel=0.0), dict(tol_abs=1e-12, tol_rel=1e-12)):
        (v, imp, info), (v_ref, imp_ref, res) = both(ctx, prob, alpha=alpha, **kw)
        assert info.iters == res.iters
        assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL


def test_chebyshev_pile_parity(ctx):
    """A settled pile (real sphere contact graph).  Through 200 fixed sweeps the bar is 1e-10.
    Run to tolerance, the loop stops at a point that rounding moves by far more than 1e-10: the
    extrapolation (w near 2 with rho clamped at 0.995) amplifies rounding-level differences
    about 100 times more than the standard loop, so the oracle's own answer shifts by ~5e-6
    when its input forces move by an ulp.  There the bar is the same iteration count and a
    difference no larger than that shift, measured here on the oracle itself."""
    pile = scenes.sphere_pile(n=5000, settle_steps=150, box=(1.5, 1.5, 3.0), seed=5)
    prob = pile.problem()
    (v, imp, info), (v_ref, imp_ref, res) = both(ctx, prob, dt=5e-4, max_iter=200, tol_abs=0.0, tol_rel=0.0)
    assert info.iters == res.iters == 200
    assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL
    kw = dict(tol_abs=1e-6, tol_rel=1e-6)
    (v, imp, info), (v_ref, imp_ref, res) = both(ctx, prob, dt=5e-4, **kw)
    pert = copy.deepcopy(prob)
    rng = np.random.default_rng(0)
    pert.fext = prob.fext + np.abs(prob.fext).max() * 4e-16 * rng.standard_normal(prob.fext.shape)
    st, v_p, imp_p, res_p = oracle.pj_solve(pert, 5e-4, oracle.params(chebyshev=1, **kw))
    assert info.iters == res.iters == res_p.iters
    assert rel(v, v_ref) <= max(TOL, rel(v_p, v_ref)) and rel(imp, imp_ref) <= max(TOL, rel(imp_p, imp_ref))


def test_chebyshev_deterministic_and_faster(ctx):
    prob = scenes.random_problem(nb=500, nc=750, seed=9)
    a = both(ctx, prob, max_iter=300, tol_abs=0.0, tol_rel=0.0)[0]
    b = both(ctx, prob, max_iter=300, tol_abs=0.0, tol_rel=0.0)[0]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # fewer sweeps than the standard loop where it converges
    prob = scenes.random_problem(**CONVERGING[0])
    (v, imp, info), _ = both(ctx, prob, tol_abs=1e-9, tol_rel=1e-9)
    p = native.default_params()
    p.tol_abs = p.tol_rel = 1e-9
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    assert ctx.pj_solve(1e-3, 1.0, p).iters > 2 * info.iters


@pytest.fixture
def rctx():
    """A context that sweeps contact ranges whatever the problem size (cdx_set_range_sweep mode 2)."""
    c = native.Context(0)
    c.set_range_sweep(2)
    yield c
    c.close()


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"nb{c['nb']}_nc{c['nc']}_s{c['seed']}")
@pytest.mark.parametrize("sweeps", [1, 2, 9, 64])
def test_chebyshev_range_sweep_fixed_sweeps(rctx, case, sweeps):
    """The Chebyshev loop on the range sweep (k_pj_range<*, *, true> + k_pjc_bodies_parts)."""
    prob = scenes.random_problem(**case)
    (v, imp, info), (v_ref, imp_ref, res) = both(rctx, prob, max_iter=sweeps, tol_abs=0.0, tol_rel=0.0)
    assert rctx.range_sweep_info()[0]
    assert info.iters == res.iters == sweeps
    assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL


@pytest.mark.parametrize("case", CONVERGING, ids=lambda c: f"nb{c['nb']}_nc{c['nc']}_s{c['seed']}")
@pytest.mark.parametrize("tol", [1e-5, 1e-9])
def test_chebyshev_range_sweep_to_tolerance(rctx, case, tol):
    prob = scenes.random_problem(**case)
    (v, imp, info), (v_ref, imp_ref, res) = both(rctx, prob, tol_abs=tol, tol_rel=tol)
    assert rctx.range_sweep_info()[0]
    assert info.iters == res.iters and res.residual <= 1
    assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL


def test_chebyshev_block_mode_keeps_the_two_passes(rctx):
    prob = scenes.random_problem(nb=50, nc=200, seed=3, p_point=0.3)
    (v, imp, info), (v_ref, imp_ref, res) = both(rctx, prob, max_iter=9, tol_abs=0.0, tol_rel=0.0, block_precond=1)
    assert not rctx.range_sweep_info()[0]
    assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL
