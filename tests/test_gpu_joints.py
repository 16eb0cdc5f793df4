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


def drop_rows(rows, keep_prob, seed):
    """A random subset of each joint's spring and damper rows: blocks of 0..6 rows, mixed along
    the chains (widths other than a full joint's 3 or 6)."""
    rng = np.random.default_rng(seed)
    ns, nd = int(rows["ns"]), int(rows["nd"])
    ks, kd = rng.random(ns) < keep_prob, rng.random(nd) < keep_prob
    sel = np.concatenate([np.flatnonzero(ks), ns + np.flatnonzero(kd)])
    out = {k: np.asarray(rows[k])[sel] for k in ("body_a", "body_b", "wa", "wb", "comp", "vsrc", "gerr")}
    out.update(ns=int(ks.sum()), nd=int(kd.sum()))
    return out


@pytest.mark.parametrize("sweeps", [1, 23])
def test_mixed_block_widths_match_oracle(ctx, sweeps):
    prob, rows = chain_case(nchains=4, links=9, seed=5)
    rows = drop_rows(rows, 0.6, seed=11)
    sizes = np.unique(np.stack([rows["body_a"], rows["body_b"]], 1), axis=0, return_counts=True)[1]
    assert len(set(sizes.tolist())) >= 4  # several block widths, not only 3 and 6
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
    for kw in (dict(accel=native.CDX_ACCEL_ANDERSON), dict(block_precond=1)):
        with pytest.raises(native.CdxError):
            ctx.pj_solve(1e-3, 1.0, params(**kw))
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    info = ctx.pj_solve(1e-3, 1.0, params())
    assert info.iters == 0 and np.allclose(ctx.velocities(), prob.v + 1e-3 * prob.minv * (prob.fext + prob.fgyro))


def chains_batch(nchains=6, links=6, seed=3, free=(), damping=0.05):
    """Capsule chains on a shared terrain, one chain per scene; the chains listed in `free` keep
    no joint rows (plain contact scenes inside the same batch)."""
    t = scenes.fractal_heightfield(n=33, extent=1.6, amp=0.03, seed=seed)
    batch, jt = scenes.capsule_chains(nchains=nchains, links=links, terrain=t, seed=seed)
    q = batch.q.copy()
    rng = np.random.default_rng(seed + 11)
    qq = q[:, 3:] + 0.15 * rng.normal(size=(len(q), 4))
    q[:, 3:] = qq / np.linalg.norm(qq, axis=1, keepdims=True)
    q[:, 2] = scenes.terrain_height(t, q[:, 0], q[:, 1]) + 0.02 - 2e-4
    v = 0.3 * rng.normal(size=(len(q), 6))
    bd = scenes.shape_bodies(batch.shape, batch.dims, batch.mass, q, v.reshape(-1))
    con = batch.contacts(q=q)
    prob = scenes.make_problem(bd, con)
    keep = ~np.isin(jt["body_a"] // links, list(free))
    jt = {k: np.asarray(x)[keep] for k, x in jt.items()}
    rows = joints.spherical_rows(q, jt, joints.joint_frames(batch.q, jt), damping)
    return prob, rows, np.asarray(batch.body_off), np.asarray(con["contact_off"])


def scene_slice(prob, rows, body_off, contact_off, s, lam_prev=None):
    """Scene s of a batch of 6-dof bodies as a standalone problem with its own joint rows."""
    b0, b1, c0, c1 = int(body_off[s]), int(body_off[s + 1]), int(contact_off[s]), int(contact_off[s + 1])

    def rebase(ids):
        ids = np.asarray(ids[c0:c1]).copy()
        ids[ids >= 0] -= b0
        return ids.astype(np.int32)

    sub = scenes.ContactProblem(dof=prob.dof[b0:b1], minv=prob.minv[6 * b0:6 * b1], v=prob.v[6 * b0:6 * b1],
                                fext=prob.fext[6 * b0:6 * b1], fgyro=prob.fgyro[6 * b0:6 * b1], q=prob.q[7 * b0:7 * b1],
                                body_a=rebase(prob.body_a), body_b=rebase(prob.body_b), mu=prob.mu[c0:c1],
                                frame=prob.frame[c0:c1], angA=prob.angA[c0:c1], angB=prob.angB[c0:c1],
                                rest=prob.rest[c0:c1], vbias=prob.vbias[c0:c1], warm=prob.warm[c0:c1])
    inside = lambda ids: (ids >= b0) & (ids < b1)
    sel = inside(rows["body_a"]) | inside(rows["body_b"])
    n = rows["ns"] + rows["nd"]
    spring = np.arange(n) < rows["ns"]
    sub_rows = {k: np.asarray(rows[k])[sel] for k in ("body_a", "body_b", "wa", "wb", "comp", "vsrc", "gerr", "joint")}
    for k in ("body_a", "body_b"):
        sub_rows[k] = np.where(sub_rows[k] >= 0, sub_rows[k] - b0, -1).astype(np.int32)
    sub_rows["ns"], sub_rows["nd"] = int((sel & spring).sum()), int((sel & ~spring).sum())
    return sub, sub_rows, (None if lam_prev is None else np.asarray(lam_prev)[sel]), sel


@pytest.mark.parametrize("launch", [(256, 1), (256, 4), (64, 1)])
@pytest.mark.parametrize("kw", [dict(max_iter=23, tol_abs=0.0, tol_rel=0.0), dict(tol_abs=1e-9, tol_rel=1e-9)],
                         ids=["fixed", "tol"])
def test_batched_chains_match_oracle(ctx, launch, kw):
    """Joint rows in the batched kernel: every scene (a chain on the terrain, or a chain without
    joints) equals the oracle's solve of that scene alone, with its own iteration count."""
    prob, rows, boff, coff = chains_batch(free=(2,))
    lam_prev = np.random.default_rng(5).normal(size=rows["ns"] + rows["nd"]) * 1e-3
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    ctx.set_joints(rows, lam_prev=lam_prev)
    ctx.set_scenes(boff, coff)
    ctx.batch_set_launch(*launch)
    it, res = ctx.batch_pj_solve(1e-3, 1.0, params(**kw))
    v, imp, lam = ctx.velocities(), ctx.impulses(), ctx.joint_impulses()
    for s in range(len(boff) - 1):
        sub, sub_rows, lp, sel = scene_slice(prob, rows, boff, coff, s, lam_prev)
        b0, b1, c0, c1 = boff[s], boff[s + 1], coff[s], coff[s + 1]
        if sub_rows["ns"] + sub_rows["nd"]:
            st, v_ref, imp_ref, lam_ref, r = oracle.pj_solve_joints(sub, sub_rows, 1e-3, oracle.params(**kw), lam_prev=lp)
            assert rel(lam[sel], lam_ref) <= TOL
        else:
            st, v_ref, imp_ref, r = oracle.pj_solve(sub, 1e-3, oracle.params(**kw))
        assert st == 0 and it[s] == r.iters
        assert rel(v[6 * b0:6 * b1], v_ref) <= TOL and rel(imp[c0:c1], imp_ref) <= TOL


def test_batched_chains_free_and_errors(ctx):
    """Free chains (no contact rows): each scene is its free solution S^-1 rhs, as the oracle's;
    a chain split between two scenes and theta = 0 are refused."""
    prob, rows = chain_case(nchains=2, links=4, lower=False)
    assert prob.nc == 0
    lam_prev = np.random.default_rng(3).normal(size=rows["ns"] + rows["nd"])
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    ctx.set_joints(rows, lam_prev=lam_prev)
    ctx.set_scenes([0, 2, 8], [0, 0, 0])  # bodies 0-3 form the first chain
    with pytest.raises(native.CdxError) as e:
        ctx.batch_pj_solve(1e-3, 1.0, params())
    assert e.value.status == native.CDX_BAD_ARG
    ctx.set_scenes([0, 4, 8], [0, 0, 0])
    with pytest.raises(native.CdxError) as e:
        ctx.batch_pj_solve(1e-3, 0.0, params())
    assert e.value.status == native.CDX_BAD_ARG
    it, res = ctx.batch_pj_solve(1e-3, 0.7, params())
    assert (it == 0).all()
    st, v_ref, _, lam_ref, r = oracle.pj_solve_joints(prob, rows, 1e-3, oracle.params(), theta=0.7, lam_prev=lam_prev)
    assert st == 0 and rel(ctx.velocities(), v_ref) <= TOL and rel(ctx.joint_impulses(), lam_ref) <= TOL


def test_long_chains_staged_in_windows(ctx):
    """60-link chains (59 joint blocks): the chain solve stages L and the row values through
    several LDS windows (grid-wide solve and batched kernel); both match the oracle."""
    t = scenes.fractal_heightfield(n=65, extent=8.0, amp=0.05, seed=8)
    batch, jt = scenes.capsule_chains(nchains=2, links=60, terrain=t, seed=8, drop=0.0005, local=True)
    rng = np.random.default_rng(8)
    v = 0.3 * rng.normal(size=(len(batch.q), 6))
    bd = scenes.shape_bodies(batch.shape, batch.dims, batch.mass, batch.q, v.reshape(-1))
    con = batch.contacts()
    prob = scenes.make_problem(bd, con)
    rows = joints.spherical_rows(batch.q, jt, joints.joint_frames(batch.q, jt), 0.05)
    assert prob.nc > 0
    kw = dict(max_iter=25, tol_abs=0.0, tol_rel=0.0)
    (vg, imp, lam, info), (v_ref, imp_ref, lam_ref, res) = run_both(ctx, prob, rows, 1e-3, **kw)
    assert info.iters == res.iters == 25
    assert rel(vg, v_ref) <= TOL and rel(imp, imp_ref) <= TOL and rel(lam, lam_ref) <= TOL
    # the batched solve starts again from the uploaded warm start (a solve leaves its impulses)
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    ctx.set_joints(rows)
    ctx.set_scenes(batch.body_off, con["contact_off"])
    ctx.batch_set_launch(256, 2)
    it, _ = ctx.batch_pj_solve(1e-3, 1.0, params(**kw))
    for s in range(batch.nscenes):
        sub, sr, lp, sel = scene_slice(prob, rows, batch.body_off, con["contact_off"], s)
        st, v1, i1, l1, r1 = oracle.pj_solve_joints(sub, sr, 1e-3, oracle.params(**kw))
        b0, b1 = batch.body_off[s], batch.body_off[s + 1]
        assert it[s] == r1.iters
        assert rel(ctx.velocities()[6 * b0:6 * b1], v1) <= TOL and rel(ctx.joint_impulses()[sel], l1) <= TOL


def test_set_joints_reuses_tables_only_for_the_same_bodies(ctx):
    """cdx_set_joints keeps the chain tables of the last call and only refreshes the row values
    when the row bodies are the same: new values on the same joints, then the same rows listed in
    another order (a new block order), then new values again -- each solve matches the oracle."""
    prob, rows = chain_case(nchains=3, links=6, seed=1)
    rng = np.random.default_rng(6)
    perm = rng.permutation(rows["ns"])  # reorder the spring rows, dampers unchanged
    n = rows["ns"] + rows["nd"]
    order = np.concatenate([perm, np.arange(rows["ns"], n)])
    permuted = {k: (np.asarray(v)[order] if isinstance(v, np.ndarray) and len(v) == n else v) for k, v in rows.items()}
    scaled = dict(rows, gerr=np.asarray(rows["gerr"]) * 1.5, comp=np.asarray(rows["comp"]) * 0.7)
    for r in (rows, scaled, permuted, dict(permuted, gerr=np.asarray(permuted["gerr"]) * 0.5)):
        (v, imp, lam, info), (v_ref, imp_ref, lam_ref, res) = run_both(ctx, prob, r, 1e-3, max_iter=20, tol_abs=0.0,
                                                                      tol_rel=0.0)
        assert info.iters == res.iters == 20
        assert rel(v, v_ref) <= TOL and rel(imp, imp_ref) <= TOL and rel(lam, lam_ref) <= TOL
