"""GPU parity: the HIP path through the C ABI against the CPU oracle on the same inputs.

Bar (from the task): FP64 projected Jacobi matches impulses and velocities to <= 1e-10 relative
after the same iteration count; coloured Gauss-Seidel matches the oracle's coloured sweep to the
same bar, and is compared with sequential PGS by complementarity residual and kinetic energy.
"Relative" is max |gpu - ref| / max |ref| over the whole vector.
"""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import native, scenes

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    a = np.asarray(a, float).ravel()
    b = np.asarray(b, float).ravel()
    den = max(np.abs(b).max(initial=0.0), 1e-300)
    return np.abs(a - b).max(initial=0.0) / den


def gpu_params(**kw):
    p = native.default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def upload(ctx, prob):
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())


def gpu_pj(ctx, prob, dt, **kw):
    upload(ctx, prob)
    info = ctx.pj_solve(dt, 1.0, gpu_params(**kw))
    return ctx.velocities(), ctx.impulses(), info


CASES = [
    dict(nb=2, nc=1, seed=0, p_static=1.0),
    dict(nb=2, nc=1, seed=1, p_static=0.0, p_frictionless=1.0),
    dict(nb=20, nc=60, seed=2),
    dict(nb=50, nc=200, seed=3, p_point=0.3),
    dict(nb=200, nc=300, seed=4, rest=0.4),
    dict(nb=300, nc=450, seed=5, p_frictionless=0.0, p_static=0.3),
    dict(nb=1000, nc=1500, seed=6, warm=False),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"nb{c['nb']}_nc{c['nc']}_s{c['seed']}")
@pytest.mark.parametrize("sweeps", [1, 10, 137])
def test_pj_fixed_sweeps(ctx, case, sweeps):
    prob = scenes.random_problem(**case)
    dt = 1e-3
    st, v_ref, imp_ref, res = oracle.pj_solve(prob, dt, oracle.params(max_iter=sweeps, tol_abs=0, tol_rel=0))
    assert st == 0 and res.iters == sweeps
    v, imp, info = gpu_pj(ctx, prob, dt, max_iter=sweeps, tol_abs=0.0, tol_rel=0.0)
    assert info.iters == sweeps
    assert info.nrows == res.nrows
    assert rel(v, v_ref) <= TOL, rel(v, v_ref)
    assert rel(imp, imp_ref) <= TOL, rel(imp, imp_ref)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("tol", [1e-4, 1e-8])
def test_pj_to_tolerance_same_iterations(ctx, seed, tol):
    prob = scenes.random_problem(nb=80, nc=120, seed=10 + seed)
    dt = 1e-3
    st, v_ref, imp_ref, res = oracle.pj_solve(prob, dt, oracle.params(tol_abs=tol, tol_rel=tol))
    v, imp, info = gpu_pj(ctx, prob, dt, tol_abs=tol, tol_rel=tol)
    assert info.iters == res.iters
    assert info.residual == pytest.approx(res.residual, rel=1e-6)
    assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL


def test_pj_warmstart_off_and_alpha(ctx):
    prob = scenes.random_problem(nb=60, nc=90, seed=21)
    for alpha, warm in ((0.7, 0), (0.1, 1)):
        st, v_ref, imp_ref, res = oracle.pj_solve(prob, 1e-3, oracle.params(alpha=alpha, warmstart=warm, max_iter=50,
                                                                            tol_abs=0, tol_rel=0))
        v, imp, info = gpu_pj(ctx, prob, 1e-3, alpha=alpha, warmstart=warm, max_iter=50, tol_abs=0.0, tol_rel=0.0)
        assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL


def test_pj_no_contacts_and_static_static(ctx):
    prob = scenes.random_problem(nb=5, nc=4, seed=2)
    prob.body_a[:] = -1
    prob.body_b[:] = -1  # static-static: no rows
    st, v_ref, imp_ref, res = oracle.pj_solve(prob, 1e-3, oracle.params())
    v, imp, info = gpu_pj(ctx, prob, 1e-3)
    assert info.nrows == 0 and info.iters == 0
    assert rel(v, v_ref) <= 1e-14
    assert (imp == 0).all()


def test_error_statuses(ctx):
    prob = scenes.random_problem(nb=4, nc=6, seed=1)
    bad = prob.copy()
    bad.minv[0:3] = 0.0
    ctx.set_bodies(bad.dof, bad.minv, bad.v, bad.fext, bad.fgyro, bad.q)
    ctx.set_contacts(*bad.contact_arrays())
    with pytest.raises(native.CdxError) as e:
        ctx.pj_solve(1e-3, 1.0, gpu_params())
    assert e.value.status == native.CDX_NOT_FACTORIZED
    upload(ctx, prob)
    bad = prob.copy()
    bad.body_b[0] = 99
    with pytest.raises(native.CdxError) as e:
        ctx.set_contacts(*bad.contact_arrays())
    assert e.value.status == native.CDX_BAD_ARG
    with pytest.raises(native.CdxError) as e:
        ctx.gs_colored_solve(1e-3, 1.0, gpu_params(alpha=1e6, max_iter=5000, tol_abs=1e-12, tol_rel=1e-12))
    assert e.value.status == native.CDX_DIVERGED


@pytest.mark.parametrize("case", CASES[2:], ids=lambda c: f"nb{c['nb']}_nc{c['nc']}_s{c['seed']}")
def test_colored_gs_matches_oracle(ctx, case):
    prob = scenes.random_problem(**case)
    dt = 1e-3
    # fixed sweep count: a tolerance that cannot be met (zero would make 0/0 = NaN, which GS
    # treats as divergence, projected_gauss_seidel.cpp:133-138)
    for sweeps, tol in ((7, 1e-30), (200000, 1e-7)):
        st, v_ref, imp_ref, res = oracle.colored_gs_solve(prob, dt, oracle.params(max_iter=sweeps, tol_abs=tol,
                                                                                  tol_rel=tol))
        assert st == 0
        upload(ctx, prob)
        info = ctx.gs_colored_solve(dt, 1.0, gpu_params(max_iter=sweeps, tol_abs=tol, tol_rel=tol))
        assert info.ncolors == res.ncolors and info.iters == res.iters
        assert rel(ctx.velocities(), v_ref) <= TOL
        assert rel(ctx.impulses(), imp_ref) <= TOL


def test_colored_gs_context_reuse_and_static_pairs(ctx):
    """The colour-ordered contact copy is rebuilt for every solve: one context solving problems of
    different sizes in turn (its buffers grow and are reused) gives, bit for bit, what a fresh
    context gives; contacts with two static sides are left out of the colouring and keep their
    impulses' slots untouched in contact order."""
    dt = 1e-3
    big = scenes.random_problem(nb=400, nc=900, seed=21)
    small = scenes.random_problem(nb=30, nc=50, seed=22, p_frictionless=0.5)
    both_static = scenes.random_problem(nb=60, nc=120, seed=23)
    both_static.body_a[::7] = -1
    both_static.body_b[::7] = -1
    prm = dict(max_iter=40, tol_abs=1e-30, tol_rel=1e-30)
    for prob in (big, small, both_static, big):
        upload(ctx, prob)
        info = ctx.gs_colored_solve(dt, 1.0, gpu_params(**prm))
        v, imp = ctx.velocities(), ctx.impulses()
        fresh = native.Context(0)
        upload(fresh, prob)
        finfo = fresh.gs_colored_solve(dt, 1.0, gpu_params(**prm))
        assert info.iters == finfo.iters and info.ncolors == finfo.ncolors
        assert np.array_equal(v, fresh.velocities()) and np.array_equal(imp, fresh.impulses())
        st, v_ref, imp_ref, res = oracle.colored_gs_solve(prob, dt, oracle.params(**prm))
        assert st == 0 and res.iters == info.iters
        assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL


def kinetic_energy(prob, v):
    return 0.5 * np.sum(v * v / prob.minv)


def test_colored_gs_vs_sequential_pgs_residual_and_energy(ctx):
    """Sweep order differs from the reference PGS, so compare by complementarity residual and
    kinetic energy at tight tolerance on a sphere pile."""
    pile = scenes.sphere_pile(n=3000, settle_steps=120, box=(1.0, 1.0, 2.0), seed=3)
    prob = pile.problem()
    dt = 5e-4
    tight = dict(tol_abs=1e-9, tol_rel=1e-9)
    st, v_pgs, imp_pgs, res = oracle.pgs_solve(prob, dt, oracle.params(**tight))
    assert st == 0
    upload(ctx, prob)
    info = ctx.gs_colored_solve(dt, 1.0, gpu_params(**tight))
    v, imp = ctx.velocities(), ctx.impulses()
    cr_pgs = oracle.complementarity(prob, v_pgs, imp_pgs * [1, -1, -1])
    cr_gpu = oracle.complementarity(prob, v, imp * [1, -1, -1])
    assert cr_gpu < 1e-4 and cr_pgs < 1e-4
    ke_ref = kinetic_energy(prob, v_pgs)
    assert kinetic_energy(prob, v) == pytest.approx(ke_ref, rel=1e-3, abs=1e-9)


def test_pj_pile_parity(ctx):
    """A settled pile (real sphere contact graph) at 200 fixed sweeps."""
    pile = scenes.sphere_pile(n=5000, settle_steps=150, box=(1.5, 1.5, 3.0), seed=5)
    prob = pile.problem()
    dt = 5e-4
    st, v_ref, imp_ref, res = oracle.pj_solve(prob, dt, oracle.params(max_iter=200, tol_abs=0, tol_rel=0))
    v, imp, info = gpu_pj(ctx, prob, dt, max_iter=200, tol_abs=0.0, tol_rel=0.0)
    assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL


def test_pj_deterministic(ctx):
    prob = scenes.random_problem(nb=500, nc=750, seed=9)
    a = gpu_pj(ctx, prob, 1e-3, max_iter=64, tol_abs=0.0, tol_rel=0.0)
    b = gpu_pj(ctx, prob, 1e-3, max_iter=64, tol_abs=0.0, tol_rel=0.0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_moreau_update_matches_oracle(ctx):
    prob = scenes.random_problem(nb=40, nc=100, seed=4, p_point=0.25)
    upload(ctx, prob)
    ctx.pj_solve(1e-3, 1.0, gpu_params(max_iter=20, tol_abs=0.0, tol_rel=0.0))
    v = ctx.velocities()
    ctx.moreau_update(1e-3, 1.0)
    q = ctx.positions()
    q_ref = oracle.position_update(prob.dof, prob.q, v, 1e-3)
    assert rel(q, q_ref) <= 1e-14
    # theta = 0.5: pre-drift with v_n, post-drift with v_{n+1}
    upload(ctx, prob)
    ctx.moreau_predrift(1e-3, 0.5)
    np.testing.assert_allclose(ctx.positions(), oracle.position_update(prob.dof, prob.q, prob.v, 0.5e-3), rtol=0,
                               atol=1e-15)


@pytest.mark.parametrize("seed", range(3))
def test_pj_nesterov_matches_oracle(ctx, seed):
    """Nesterov loop (projected_jacobi.cpp:379-449): residual every sweep, restarts on residual
    growth / beta threshold / direction reversal, momentum disabled after the restart limit."""
    prob = scenes.random_problem(nb=120, nc=180, seed=30 + seed)
    dt = 1e-3
    for kw in (dict(max_iter=40, tol_abs=0.0, tol_rel=0.0), dict(tol_abs=1e-8, tol_rel=1e-8),
               dict(tol_abs=1e-8, tol_rel=1e-8, nesterov_restart_limit=3), dict(tol_abs=1e-8, tol_rel=1e-8,
                                                                                 nesterov_beta_threshold=0.5)):
        okw = {k: v for k, v in kw.items()}
        st, v_ref, imp_ref, res = oracle.pj_solve(prob, dt, oracle.params(nesterov=1, **okw))
        v, imp, info = gpu_pj(ctx, prob, dt, accel=native.CDX_ACCEL_NESTEROV, **kw)
        assert info.iters == res.iters
        assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL
