"""GPU parity of the Nesterov and Chebyshev loops with joint rows (nesterov_loop, chebyshev_loop
and estimateSpectralRadius, projected_jacobi.cpp:155-228,379-449, over x = [v; Lambda] with the
extended S of PjAssembler) against the oracle's restatement: velocities, contact impulses and
joint impulses to <= 1e-10 relative after the same sweeps, the same iteration count to tolerance,
the same spectral radius; chains on terrain with contacts, free chains, a chain without joint
rows beside jointed ones, the restart rules, and the refusals that remain."""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import native
from test_gpu_joints import chain_case, params, rel, run_both

pytestmark = pytest.mark.gpu

TOL = 1e-10
ACCEL = {"nesterov": native.CDX_ACCEL_NESTEROV, "chebyshev": native.CDX_ACCEL_CHEBYSHEV}


def both(ctx, prob, rows, kind, **kw):
    oracle_kw = dict(nesterov=int(kind == "nesterov"), chebyshev=int(kind == "chebyshev"))
    st, v_ref, imp_ref, lam_ref, res = oracle.pj_solve_joints(prob, rows, 1e-3, oracle.params(**oracle_kw, **kw))
    assert st == 0
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    ctx.set_joints(rows)
    info = ctx.pj_solve(1e-3, 1.0, params(accel=ACCEL[kind], **kw))
    return (ctx.velocities(), ctx.impulses(), ctx.joint_impulses(), info), (v_ref, imp_ref, lam_ref, res)


def check(gpu, ref, iters=None):
    (v, imp, lam, info), (v_ref, imp_ref, lam_ref, res) = gpu, ref
    assert info.iters == res.iters
    if iters is not None:
        assert info.iters == iters
    assert rel(v, v_ref) <= TOL, rel(v, v_ref)
    assert rel(imp, imp_ref) <= TOL, rel(imp, imp_ref)
    assert rel(lam, lam_ref) <= TOL, rel(lam, lam_ref)


CASES = [dict(seed=0), dict(seed=1, damping=0.4), dict(seed=2, nchains=4, links=5, drop_chain=1)]


@pytest.mark.parametrize("kind", ["nesterov", "chebyshev"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "_".join(f"{k}{v}" for k, v in c.items()))
def test_fixed_sweeps_and_tolerance(ctx, kind, case):
    prob, rows = chain_case(**case)
    assert prob.nc > 0
    # fixed sweep counts (an unreachable tolerance), then to tolerance with the same count
    for sweeps in (1, 2, 7, 33):
        check(*both(ctx, prob, rows, kind, max_iter=sweeps, tol_abs=1e-30, tol_rel=1e-30), iters=sweeps)
    gpu, ref = both(ctx, prob, rows, kind, max_iter=20000, tol_abs=1e-8, tol_rel=1e-8)
    assert ref[3].iters < 20000
    check(gpu, ref)


@pytest.mark.parametrize("kind", ["nesterov", "chebyshev"])
def test_free_chains(ctx, kind):
    prob, rows = chain_case(nchains=2, links=6, lower=False)
    assert prob.nc == 0  # no contact rows: the free solution, no loop
    check(*both(ctx, prob, rows, kind, max_iter=50, tol_abs=1e-8, tol_rel=1e-8), iters=0)


@pytest.mark.parametrize("rules", [dict(nesterov_beta_threshold=0.3), dict(nesterov_restart_limit=2)],
                         ids=["beta_threshold", "restart_limit"])
def test_nesterov_restart_rules(ctx, rules):
    prob, rows = chain_case(seed=3)
    for sweeps, tol in ((25, 1e-30), (20000, 1e-8)):
        check(*both(ctx, prob, rows, "nesterov", max_iter=sweeps, tol_abs=tol, tol_rel=tol, **rules))


def test_the_standard_loop_still_matches_after_an_accelerated_solve(ctx):
    """The accelerated loop's state buffers are its own: a standard solve on the same context
    afterwards equals the oracle's standard solve."""
    prob, rows = chain_case(seed=4)
    both(ctx, prob, rows, "chebyshev", max_iter=40, tol_abs=1e-30, tol_rel=1e-30)
    (v, imp, lam, info), (v_ref, imp_ref, lam_ref, res) = run_both(ctx, prob, rows, 1e-3, max_iter=40, tol_abs=0.0,
                                                                  tol_rel=0.0)
    assert info.iters == res.iters == 40
    assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL and rel(lam, lam_ref) <= TOL
