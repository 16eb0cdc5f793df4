"""Gradients through joint rows: the implicit adjoint of the extended-system solve
(cdx_pj_adjoint_joints) against central finite differences of the oracle's forward solve run to
tight tolerance, for dL/d(joint damping), dL/dv_n, dL/df_ext and dL/dmu, after the grid-wide and
the batched solve."""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import joints, native, scenes
from test_gpu_joints import chain_case, chains_batch, params

pytestmark = pytest.mark.gpu

DT = 1e-3
TIGHT = dict(tol_abs=1e-14, tol_rel=1e-14, max_iter=2_000_000)


def forward(prob, rows, theta=1.0, lam_prev=None):
    st, v, imp, lam, res = oracle.pj_solve_joints(prob, rows, DT, oracle.params(**TIGHT), theta=theta, lam_prev=lam_prev)
    assert st == 0
    return v


def device_grad(ctx, prob, rows, gv, theta=1.0, lam_prev=None, batched=None):
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    ctx.set_joints(rows, lam_prev=lam_prev)
    p = params(tol_abs=1e-13, tol_rel=1e-13)
    if batched is None:
        ctx.pj_solve(DT, theta, p)
    else:
        ctx.set_scenes(*batched)
        ctx.batch_pj_solve(DT, theta, p)
    return ctx.pj_adjoint_joints(DT, p, gv, tol=1e-13)


def with_damping(rows, joint, D):
    """Rows with the damper rows of `joint` at damping D (A = 1/D)."""
    r = {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v) for k, v in rows.items()}
    sel = (r["joint"] == joint) & (np.arange(r["ns"] + r["nd"]) >= r["ns"])
    r["comp"][sel] = 1.0 / D
    return r


def fd(f, x0, h):
    return (f(x0 + h) - f(x0 - h)) / (2 * h)


def check(a, b, rtol, atol):
    assert abs(a - b) <= rtol * max(abs(a), abs(b)) + atol, (a, b)


@pytest.mark.parametrize("lower", [True, False], ids=["terrain", "free"])
def test_damping_gradient_matches_finite_differences(ctx, lower):
    prob, rows = chain_case(nchains=2, links=5, seed=1, lower=lower, damping=0.05)
    gv = np.random.default_rng(0).normal(size=prob.nv)
    out = device_grad(ctx, prob, rows, gv)
    assert out["info"].iters >= 0
    nd = rows["ns"]
    for j in (0, 3, 6):
        sel = (rows["joint"] == j) & (np.arange(rows["ns"] + rows["nd"]) >= nd)
        dA = out["comp"][sel].sum()
        D = 0.05
        g_dev = -dA / D ** 2
        g_fd = fd(lambda d: gv @ forward(prob, with_damping(rows, j, d)), D, 1e-4 * D)
        check(g_dev, g_fd, 1e-5, 1e-9 * np.abs(gv).max())


def test_velocity_force_and_friction_gradients(ctx):
    prob, rows = chain_case(nchains=2, links=5, seed=1, lower=True)
    assert prob.nc > 0
    rng = np.random.default_rng(1)
    gv = rng.normal(size=prob.nv)
    out = device_grad(ctx, prob, rows, gv)
    for i in rng.choice(prob.nv, size=6, replace=False):
        def L(x, field):
            p = prob.copy()
            getattr(p, field)[i] = x
            p.fgyro = scenes.gyro_force(p.dof, p.minv, p.v)  # f_gyro = -w x (I w) follows v
            return gv @ forward(p, rows)
        check(out["v"][i], fd(lambda x: L(x, "v"), prob.v[i], 1e-6), 1e-5, 1e-8)
        check(out["fext"][i], fd(lambda x: L(x, "fext"), prob.fext[i], 1e-4), 1e-5, 1e-10)
    for c in rng.choice(prob.nc, size=4, replace=False):
        def Lmu(x):
            p = prob.copy()
            p.mu[c] = x
            return gv @ forward(p, rows)
        # central differences with h = 1e-6 of solves converged to 1e-13 carry ~1e-7 of noise
        check(out["mu"][c], fd(Lmu, prob.mu[c], 1e-6), 1e-4, 1e-7)


def test_theta_and_spring_compliance(ctx):
    """theta < 1 adds the (1 - theta) G v term to the joint rows' right-hand side; compliant
    spring rows (C > 0) with Lambda_prev give the dL/dC of E and of the Baumgarte term."""
    prob, rows = chain_case(nchains=2, links=4, seed=2, lower=True)
    rows = dict(rows)
    rows["comp"] = rows["comp"].copy()
    rows["comp"][: rows["ns"]] = 1e-7
    lam_prev = np.random.default_rng(4).normal(size=rows["ns"] + rows["nd"]) * 1e-2
    gv = np.random.default_rng(2).normal(size=prob.nv)
    out = device_grad(ctx, prob, rows, gv, theta=0.7, lam_prev=lam_prev)
    rng = np.random.default_rng(3)
    for i in rng.choice(prob.nv, size=4, replace=False):
        def L(x):
            p = prob.copy()
            p.v[i] = x
            p.fgyro = scenes.gyro_force(p.dof, p.minv, p.v)
            return gv @ forward(p, rows, theta=0.7, lam_prev=lam_prev)
        check(out["v"][i], fd(L, prob.v[i], 1e-6), 1e-5, 1e-8)
    for r in (0, 4, rows["ns"] + 1):
        def Lc(x):
            rr = dict(rows)
            rr["comp"] = rows["comp"].copy()
            rr["comp"][r] = x
            return gv @ forward(prob, rr, theta=0.7, lam_prev=lam_prev)
        c0 = rows["comp"][r]
        check(out["comp"][r], fd(Lc, c0, 1e-3 * c0), 1e-4, 1e-9 * abs(out["comp"]).max())


def test_after_batched_solve_equals_grid_solve(ctx):
    prob, rows, boff, coff = chains_batch(nchains=4, links=5, seed=6)
    gv = np.random.default_rng(7).normal(size=prob.nv)
    a = device_grad(ctx, prob, rows, gv)
    b = device_grad(ctx, prob, rows, gv, batched=(boff, coff))
    for k in ("v", "fext", "mu", "comp"):
        np.testing.assert_allclose(b[k], a[k], rtol=1e-8, atol=1e-10 * max(np.abs(a[k]).max(), 1e-300))


def test_refusals(ctx):
    prob, rows = chain_case(nchains=1, links=3, lower=False)
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    with pytest.raises(native.CdxError):  # no joint solve yet
        ctx.pj_adjoint_joints(DT, params(), np.ones(prob.nv))
    ctx.set_joints(rows)
    ctx.pj_solve(DT, 1.0, params())
    with pytest.raises(native.CdxError) as e:  # the row-scaled contact adjoint does not take joint rows
        ctx.pj_adjoint(DT, params(), np.ones(prob.nv))
    assert e.value.status == native.CDX_NO_STATE
    out = ctx.pj_adjoint_joints(DT, params(), np.ones(prob.nv))
    assert np.isfinite(out["v"]).all() and np.isfinite(out["comp"]).all()


def test_chain_rollout_damping_gradient_vs_finite_differences():
    """Checkpointed rollout of chains dropping onto the terrain: dL/dD per joint and dL/dv0 of
    the COM loss against central differences of the same rollout with the geometry of every step
    (contacts and joint Jacobians) held at the nominal trajectory's, as the adjoint holds it."""
    from cardillocxx_amd import diff

    t = scenes.fractal_heightfield(n=33, extent=1.6, amp=0.03, seed=4)
    batch, jt = scenes.capsule_chains(nchains=2, links=5, terrain=t, seed=4, drop=0.0005, local=True)
    batch.v[:] = np.random.default_rng(9).normal(scale=0.2, size=batch.v.shape)
    r = diff.ChainRollout(batch, jt, steps=14, dt=1e-3, checkpoint_every=5, tol=1e-12, adjoint_tol=1e-13)
    target = np.random.default_rng(0).normal(scale=0.1, size=(batch.nscenes, 3))
    D0 = np.full(len(jt["body_a"]), 0.05)
    v0 = batch.v.reshape(-1).copy()
    pos = r.forward(v0, D0)
    L, gpos, _ = diff.com_loss(pos, batch.mass, batch.body_off, target)
    g = r.backward(gpos)
    assert sum(len(c["body_a"]) for c in r.contacts_log) > 0
    r.frozen, r.frozen_rows = r.contacts_log, r.rows_log

    def loss(v, D):
        return diff.com_loss(r.forward(v, D), batch.mass, batch.body_off, target)[0]

    rng = np.random.default_rng(1)
    for _ in range(2):
        d = rng.normal(size=D0.shape)
        eps = 1e-5
        fdv = (loss(v0, D0 + eps * d * D0) - loss(v0, D0 - eps * d * D0)) / (2 * eps)
        assert float(g.damping @ (d * D0)) == pytest.approx(fdv, rel=1e-4, abs=1e-12)
        dv = rng.normal(size=v0.shape)
        fdv = (loss(v0 + 1e-6 * dv, D0) - loss(v0 - 1e-6 * dv, D0)) / 2e-6
        assert float(g.v0.reshape(-1) @ dv) == pytest.approx(fdv, rel=1e-4, abs=1e-12)


def test_chain_rollout_stored_steps_equal_checkpoint_recompute():
    """The backward from the forward's stored step states gives the gradients of the checkpointed
    backward that recomputes each segment, bitwise (the recomputed steps are the same solves)."""
    from cardillocxx_amd import diff

    t = scenes.fractal_heightfield(n=33, extent=1.6, amp=0.03, seed=4)
    batch, jt = scenes.capsule_chains(nchains=2, links=5, terrain=t, seed=4, drop=0.0005, local=True)
    batch.v[:] = np.random.default_rng(9).normal(scale=0.2, size=batch.v.shape)
    target = np.random.default_rng(0).normal(scale=0.1, size=(batch.nscenes, 3))
    D0 = np.full(len(jt["body_a"]), 0.05)
    grads = []
    for store in (True, False):
        r = diff.ChainRollout(batch, jt, steps=12, dt=1e-3, checkpoint_every=5, tol=1e-12, adjoint_tol=1e-13,
                              store_steps=store)
        pos = r.forward(batch.v.reshape(-1).copy(), D0)
        grads.append(r.backward(diff.com_loss(pos, batch.mass, batch.body_off, target)[1]))
        assert (r.records is not None) == store
    np.testing.assert_array_equal(grads[0].damping, grads[1].damping)
    np.testing.assert_array_equal(grads[0].v0, grads[1].v0)
    np.testing.assert_array_equal(grads[0].mu, grads[1].mu)


def test_adjoint_graph_cycles_equal_direct_launches(ctx):
    """The GMRES restart cycle replayed as a graph (several restarts, with the host's least-squares
    work in between) gives the directly launched adjoint bitwise; profiling launches every kernel
    directly."""
    prob, rows = chain_case(nchains=12, links=14, seed=2, lower=True, damping=0.05)
    gv = np.random.default_rng(4).normal(size=prob.nv)
    outs = []
    for prof in (False, True, False):
        ctx.set_profiling(prof)
        outs.append(device_grad(ctx, prob, rows, gv))
    ctx.set_profiling(False)
    assert outs[0]["info"].iters > 62  # the cycle graph replayed at least once
    for o in outs[1:]:
        assert o["info"].iters == outs[0]["info"].iters
        for key in ("v", "fext", "mu", "comp"):
            np.testing.assert_array_equal(o[key], outs[0][key])
