"""GPU parity of the Nesterov-accelerated coloured Gauss-Seidel (the pj_nesterov branch of
ProjectedGaussSeidelSolver::solve, projected_gauss_seidel.cpp:196-262, over the coloured sweep of
condensed_solver.cpp:220-251) against the oracle's restatement in the same colour order:
velocities and impulses to <= 1e-10 relative after the same sweeps and the same iteration count to
tolerance, with the restart rules exercised (beta threshold, restart limit), graph replay equal to
direct launches bit for bit, divergence on a non-finite residual, and the adapter mapping of
pj.nesterov.  "Relative" is max |gpu - ref| / max |ref| over the whole vector."""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import native, scenes, solver

pytestmark = pytest.mark.gpu

TOL = 1e-10
DT = 1e-3


def rel(a, b):
    a = np.asarray(a, float).ravel()
    b = np.asarray(b, float).ravel()
    return np.abs(a - b).max(initial=0.0) / max(np.abs(b).max(initial=0.0), 1e-300)


def params(**kw):
    p = native.default_params()
    p.accel = native.CDX_ACCEL_NESTEROV
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def gpu(ctx, prob, **kw):
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    info = ctx.gs_colored_solve(DT, 1.0, params(**kw))
    return ctx.velocities(), ctx.impulses(), info


def ref(prob, **kw):
    st, v, imp, res = oracle.colored_gs_solve(prob, DT, oracle.params(nesterov=1, **kw))
    assert st == 0
    return v, imp, res


CASES = [
    dict(nb=20, nc=60, seed=2),
    dict(nb=50, nc=200, seed=3, p_point=0.3),
    dict(nb=200, nc=300, seed=4, rest=0.4),
    dict(nb=300, nc=450, seed=5, p_frictionless=0.0, p_static=0.3),
    dict(nb=1000, nc=1500, seed=6, warm=False),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"nb{c['nb']}_nc{c['nc']}_s{c['seed']}")
def test_matches_oracle(ctx, case):
    prob = scenes.random_problem(**case)
    # fixed sweep counts (a tolerance that cannot be met: zero would make 0/0, a divergence), then
    # to tolerance with the same iteration count
    for sweeps, tol in ((1, 1e-30), (6, 1e-30), (45, 1e-30), (200000, 1e-7)):
        v_ref, imp_ref, res = ref(prob, max_iter=sweeps, tol_abs=tol, tol_rel=tol)
        v, imp, info = gpu(ctx, prob, max_iter=sweeps, tol_abs=tol, tol_rel=tol)
        assert info.ncolors == res.ncolors and info.iters == res.iters
        assert rel(v, v_ref) <= TOL, (sweeps, rel(v, v_ref))
        assert rel(imp, imp_ref) <= TOL, (sweeps, rel(imp, imp_ref))


@pytest.mark.parametrize("rules", [dict(nesterov_beta_threshold=0.4), dict(nesterov_restart_limit=2)],
                         ids=["beta_threshold", "restart_limit"])
def test_restart_rules(ctx, rules):
    prob = scenes.random_problem(nb=300, nc=600, seed=8)
    for sweeps, tol in ((30, 1e-30), (200000, 1e-7)):
        v_ref, imp_ref, res = ref(prob, max_iter=sweeps, tol_abs=tol, tol_rel=tol, **rules)
        v, imp, info = gpu(ctx, prob, max_iter=sweeps, tol_abs=tol, tol_rel=tol, **rules)
        assert info.iters == res.iters
        assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL


def test_pile_to_tolerance(ctx):
    pile = scenes.sphere_pile(n=3000, settle_steps=120, box=(1.2, 1.2, 2.5), seed=3)
    prob = pile.problem()
    v_ref, imp_ref, res = ref(prob, max_iter=20000, tol_abs=1e-8, tol_rel=1e-8)
    v, imp, info = gpu(ctx, prob, max_iter=20000, tol_abs=1e-8, tol_rel=1e-8)
    assert info.iters == res.iters < 20000
    assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL


def test_graphs_equal_direct_launches(ctx):
    prob = scenes.random_problem(nb=500, nc=1200, seed=11)
    a = gpu(ctx, prob, max_iter=40, tol_abs=1e-30, tol_rel=1e-30)
    ctx.set_graphs(False)
    try:
        b = gpu(ctx, prob, max_iter=40, tol_abs=1e-30, tol_rel=1e-30)
    finally:
        ctx.set_graphs(True)
    assert a[2].iters == b[2].iters == 40
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_nonfinite_residual_is_divergence(ctx):
    prob = scenes.random_problem(nb=40, nc=80, seed=12)
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    info = ctx.gs_colored_solve(DT, 1.0, params(max_iter=50, tol_abs=0.0, tol_rel=0.0), raise_on_divergence=False)
    assert info.iters == 1
    with pytest.raises(native.CdxError) as e:
        ctx.gs_colored_solve(DT, 1.0, params(max_iter=50, tol_abs=0.0, tol_rel=0.0))
    assert e.value.status == native.CDX_DIVERGED


def test_refuses_other_accelerations(ctx):
    prob = scenes.random_problem(nb=20, nc=40, seed=13)
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    for mode in (native.CDX_ACCEL_CHEBYSHEV, native.CDX_ACCEL_ANDERSON):
        with pytest.raises(native.CdxError):
            ctx.gs_colored_solve(DT, 1.0, params(accel=mode))


def test_adapter_maps_pj_nesterov():
    """The Python mirror of SolverBase passes pj.nesterov to the coloured GS (and nothing else:
    pj.chebyshev is ignored by the reference's PGS)."""
    prob = scenes.random_problem(nb=100, nc=250, seed=14)
    # the oracle first: the adapter writes the solve's impulses back as the next warm start
    v_ref, imp_ref, res = ref(prob, max_iter=30, tol_abs=1e-30, tol_rel=1e-30)
    cfg = solver.Config(friction_enable=True, pj_nesterov=True, pj_chebyshev=True, pj_max_iterations=30,
                        pj_tol_abs=1e-30, pj_tol_rel=1e-30)
    s = solver.ColoredGaussSeidelSolver(prob, cfg)
    v = s.solve(DT, 1.0)
    assert s.last_iterations() == res.iters == 30
    assert rel(v, v_ref) <= TOL
    s.ctx.close()
