"""GPU parity of the extended-system projected Jacobi (joint rows in S, solved per chain on the
device) against the oracle's dense-LU restatement: velocities, contact impulses and joint
impulses to <= 1e-10 relative after the same sweeps, the same iteration count to tolerance, and
the error paths of cdx_set_joints."""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import joints, native, scenes

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    a = np.asarray(a, float).ravel()
    b = np.asarray(b, float).ravel()
    return np.abs(a - b).max(initial=0.0) / max(np.abs(b).max(initial=0.0), 1e-300)


def params(**kw):
    p = native.default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def chain_case(nchains=3, links=6, seed=0, lower=True, damping=0.05, drop_chain=None):
    t = scenes.fractal_heightfield(n=33, extent=1.6, amp=0.03, seed=seed)
    batch, jt = scenes.capsule_chains(nchains=nchains, links=links, terrain=t, seed=seed)
    q = batch.q.copy()
    rng = np.random.default_rng(seed + 7)
    qq = q[:, 3:] + 0.15 * rng.normal(size=(len(q), 4))
    q[:, 3:] = qq / np.linalg.norm(qq, axis=1, keepdims=True)
    if lower:
        q[:, 2] = scenes.terrain_height(t, q[:, 0], q[:, 1]) + 0.02 - 2e-4
    else:
        q[:, 2] += 1.0
    v = 0.3 * rng.normal(size=(len(q), 6))
    bd = scenes.shape_bodies(batch.shape, batch.dims, batch.mass, q, v.reshape(-1))
    con = batch.contacts(q=q)
    prob = scenes.make_problem(bd, con)
    if drop_chain is not None:  # leave one chain's links free (no joint rows)
        keep = (jt["body_a"] // links) != drop_chain
        jt = {k: np.asarray(x)[keep] for k, x in jt.items()}
    rows = joints.spherical_rows(q, jt, joints.joint_frames(batch.q, jt), damping)
    return prob, rows


def run_both(ctx, prob, rows, dt, theta=1.0, lam_prev=None, **kw):
    st, v_ref, imp_ref, lam_ref, res = oracle.pj_solve_joints(prob, rows, dt, oracle.params(**kw), theta=theta,
                                                              lam_prev=lam_prev)
    assert st == 0
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    ctx.set_joints(rows, lam_prev=lam_prev)
    info = ctx.pj_solve(dt, theta, params(**kw))
    return (ctx.velocities(), ctx.impulses(), ctx.joint_impulses(), info), (v_ref, imp_ref, lam_ref, res)


@pytest.mark.parametrize("theta", [1.0, 0.6])
def test_free_chains_match_oracle(ctx, theta):
    prob, rows = chain_case(lower=False)
    assert prob.nc == 0
    lam_prev = np.random.default_rng(3).normal(size=rows["ns"] + rows["nd"])
    (v, imp, lam, info), (v_ref, _, lam_ref, res) = run_both(ctx, prob, rows, 1e-3, theta=theta, lam_prev=lam_prev)
    assert info.iters == 0 == res.iters
    assert rel(v, v_ref) <= TOL and rel(lam, lam_ref) <= TOL


@pytest.mark.parametrize("sweeps", [1, 10, 57])
def test_chains_on_terrain_fixed_sweeps(ctx, sweeps):
    prob, rows = chain_case(seed=1)
    assert prob.nc > 0
    (v, imp, lam, info), (v_ref, imp_ref, lam_ref, res) = run_both(ctx, prob, rows, 1e-3, max_iter=sweeps, tol_abs=0.0,
                                                                  tol_rel=0.0)
    assert info.iters == res.iters == sweeps
    assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL and rel(lam, lam_ref) <= TOL


@pytest.mark.parametrize("tol", [1e-5, 1e-9])
def test_chains_on_terrain_to_tolerance(ctx, tol):
    prob, rows = chain_case(seed=2, drop_chain=1)
    (v, imp, lam, info), (v_ref, imp_ref, lam_ref, res) = run_both(ctx, prob, rows, 1e-3, tol_abs=tol, tol_rel=tol)
    assert info.iters == res.iters > 1
    assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL and rel(lam, lam_ref) <= TOL


def test_chain_rollout_matches_oracle(ctx):
    """Twenty Moreau steps of chains falling onto the terrain: joint rows, contacts and warm
    starts regenerated every step, Lambda carried over, each step against the oracle."""
    t = scenes.fractal_heightfield(n=33, extent=1.6, amp=0.03, seed=4)
    batch, jt = scenes.capsule_chains(nchains=2, links=5, terrain=t, seed=4, drop=0.01)
    # lower each straight chain until its lowest capsule touches the terrain
    for s in range(batch.nscenes):
        b0, b1 = batch.body_off[s], batch.body_off[s + 1]
        gap = batch.q[b0:b1, 2] - scenes.terrain_height(t, batch.q[b0:b1, 0], batch.q[b0:b1, 1]) - batch.dims[b0:b1, 0]
        batch.q[b0:b1, 2] -= gap.min() + 1e-4
    frames = joints.joint_frames(batch.q, jt)
    dt, q, v = 1e-3, batch.q.copy().reshape(-1), batch.v.copy().reshape(-1)
    lam_prev, prev = None, None
    worst = 0.0
    ncon = 0
    for step in range(20):
        bd = scenes.shape_bodies(batch.shape, batch.dims, batch.mass, q, v)
        con = batch.contacts(q=q.reshape(-1, 7))
        con["warm"] = scenes.match_warmstart(prev, con) if prev is not None and len(con["body_a"]) else con["warm"]
        prob = scenes.make_problem(bd, con)
        rows = joints.spherical_rows(q, jt, frames, 0.05)
        (vg, imp, lam, info), (v_ref, imp_ref, lam_ref, res) = run_both(ctx, prob, rows, dt, lam_prev=lam_prev,
                                                                        tol_abs=1e-9, tol_rel=1e-9)
        assert info.iters == res.iters
        worst = max(worst, rel(vg, v_ref), rel(lam, lam_ref), rel(imp, imp_ref) if prob.nc else 0.0)
        ncon = max(ncon, prob.nc)
        q = oracle.position_update(bd["dof"], q, v_ref, dt)
        v, lam_prev = v_ref, lam_ref
        prev = dict(body_a=con["body_a"], body_b=con["body_b"], point=con["point"], imp=imp_ref)
    assert ncon > 0 and worst <= TOL


def test_joint_errors(ctx):
    prob, rows = chain_case(nchains=1, links=4, lower=False)
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    # a branch: body 1 in three blocks
    bad = dict(rows)
    for k in ("body_a", "body_b"):
        bad[k] = rows[k].copy()
    bad["body_a"][0:3], bad["body_b"][0:3] = 1, 3
    bad["body_a"][3:6], bad["body_b"][3:6] = 1, 2
    bad["body_a"][6:9], bad["body_b"][6:9] = 0, 1
    with pytest.raises(native.CdxError) as e:
        ctx.set_joints(bad)
    assert e.value.status == native.CDX_BAD_ARG
    # a loop: 0-1, 1-2, 2-0
    loop = {k: (np.asarray(x)[:9].copy() if k not in ("ns", "nd") else x) for k, x in rows.items()}
    loop["ns"], loop["nd"] = 9, 0
    loop["body_a"][:] = np.repeat([0, 1, 2], 3)
    loop["body_b"][:] = np.repeat([1, 2, 0], 3)
    with pytest.raises(native.CdxError) as e:
        ctx.set_joints(loop)
    assert e.value.status == native.CDX_BAD_ARG
    # the other solvers refuse joint rows; set_bodies clears them
    ctx.set_joints(rows)
    with pytest.raises(native.CdxError):
        ctx.gs_colored_solve(1e-3, 1.0, params())
    for accel in (native.CDX_ACCEL_NESTEROV, native.CDX_ACCEL_CHEBYSHEV):
        with pytest.raises(native.CdxError):
            ctx.pj_solve(1e-3, 1.0, params(accel=accel))
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    info = ctx.pj_solve(1e-3, 1.0, params())
    assert info.iters == 0 and np.allclose(ctx.velocities(), prob.v + 1e-3 * prob.minv * (prob.fext + prob.fgyro))
