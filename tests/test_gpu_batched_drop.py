"""Batched drop of spheres and boxes through the one-launch batched solve, against the oracle.

The workload is the batched benchmark's own scene builder (scenes.batched_drop: bodies per scene
log-uniform in [8, 512], spheres and boxes 50/50, box edges U[0.05, 0.2] m, dropped into a unit
box), on 32 scenes.  The scenes are first settled on the device so that they carry box face
manifolds (up to 4 points per box-box or box-wall pair) and box-sphere contacts; then every
Moreau step generates contacts on the CPU, matches the previous step's impulses into the warm
start (ContactTracker's greedy rule, contact_tracker.cpp:43-72), runs cdx_batch_pj_solve, and
compares every scene with the oracle's standalone solve of that scene (projected_jacobi.cpp:451-500)
on the same inputs: the same iteration count, velocities and impulses within 1e-10 relative.
The gyroscopic force of the non-isotropic boxes is computed on the device (cdx_set_bodies with
fgyro = NULL, then after each cdx_moreau_update) and by the oracle from the same velocities
(world.cpp:152-173), so the test covers it as the solve consumes it.
"""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import native, scenes

pytestmark = pytest.mark.gpu

TOL = 1e-10
LDS_BUDGET = 60 * 1024  # the batched solve keeps scenes in LDS when they fit in this many bytes


def rel(a, b):
    den = max(np.abs(b).max(initial=0.0), 1e-300)
    return np.abs(np.asarray(a) - np.asarray(b)).max(initial=0.0) / den


def scene_problem(bd, con, v, q, fg, b0, b1, c0, c1):
    def rebase(ids):
        ids = np.asarray(ids[c0:c1]).copy()
        ids[ids >= 0] -= b0
        return ids.astype(np.int32)

    sl = slice(6 * b0, 6 * b1)
    return scenes.ContactProblem(
        dof=bd["dof"][b0:b1], minv=bd["minv"][sl], v=v[sl], fext=bd["fext"][sl], fgyro=fg[sl], q=q[7 * b0:7 * b1],
        body_a=rebase(con["body_a"]), body_b=rebase(con["body_b"]), mu=con["mu"][c0:c1], frame=con["frame"][c0:c1],
        angA=con["angA"][c0:c1], angB=con["angB"][c0:c1], rest=con["rest"][c0:c1], vbias=con["vbias"][c0:c1],
        warm=con["warm"][c0:c1])


def lds_bytes(body_off, contact_off):
    nbm = int(np.diff(body_off).max())
    ncm = int(np.diff(contact_off).max())
    return 8 * (6 * nbm + 9 * ncm) + 4 * (nbm + 1 + 2 * ncm)


def device_steps(ctx, b, bd, q, v, prev, steps, dt, prm, check=None):
    """Moreau steps of the batch on the device, contacts and warm start on the CPU."""
    ctx.set_bodies(bd["dof"], bd["minv"], v, bd["fext"], None, q)
    for s in range(steps):
        con = b.contacts(q=q.reshape(-1, 7))
        nc = len(con["body_a"])
        con["warm"] = scenes.match_warmstart(prev, con) if prev is not None and nc else np.zeros((nc, 3))
        ctx.set_contacts(*(con[k] for k in ("body_a", "body_b", "mu", "frame", "angA", "angB", "rest", "vbias", "warm")))
        ctx.set_scenes(b.body_off, con["contact_off"])
        iters, res = ctx.batch_pj_solve(dt, 1.0, prm)
        vn = ctx.velocities()
        imp = ctx.impulses()
        if check is not None:
            check(s, con, v, q, iters, vn, imp)
        ctx.moreau_update(dt, 1.0)
        qn = ctx.positions()
        if check is not None:
            q_ref = oracle.position_update(bd["dof"], q, vn, dt)
            assert rel(qn, q_ref) <= 1e-14
        q, v = qn, vn
        prev = dict(body_a=con["body_a"], body_b=con["body_b"], point=con["point"], imp=imp)
    return q, v, prev, con


@pytest.fixture(scope="module")
def settled(ctx):
    """32 scenes after 0.4 s of fall at h = 2e-3 s (input generation: loose tolerance)."""
    b = scenes.batched_drop(nscenes=32, seed=3)
    bd = b.bodies()
    prm = native.default_params()
    prm.tol_abs = prm.tol_rel = 1e-4
    prm.max_iter = 2000
    ctx.batch_set_launch(256, 1)
    q, v, prev, con = device_steps(ctx, b, bd, bd["q"].copy(), bd["v"].copy(), None, 200, 2e-3, prm)
    return b, bd, q, v, prev


@pytest.mark.parametrize("launch", [(256, 1), (512, 4)])
def test_batched_drop_matches_oracle_per_scene(ctx, settled, launch):
    b, bd, q, v, prev = settled
    dt = 1e-3
    prm = native.default_params()  # reference defaults: tol 1e-5, warm start on
    prm.max_iter = 20000
    ctx.batch_set_launch(*launch)
    seen = dict(contacts=0, max_per_scene=0, manifolds=0, lds=[], worst_v=0.0, worst_p=0.0, sweeps=0)

    def check(s, con, v_n, q_n, iters, vn, imp):
        coff = con["contact_off"]
        fg = oracle.gyro_force(bd["dof"], bd["minv"], v_n)
        seen["lds"].append(lds_bytes(b.body_off, coff) <= LDS_BUDGET)
        for k in range(b.nscenes):
            b0, b1 = int(b.body_off[k]), int(b.body_off[k + 1])
            c0, c1 = int(coff[k]), int(coff[k + 1])
            prob = scene_problem(bd, con, v_n, q_n, fg, b0, b1, c0, c1)
            st, v_ref, imp_ref, r = oracle.pj_solve(prob, dt, oracle.params(max_iter=prm.max_iter))
            assert st == 0
            assert iters[k] == r.iters, (s, k)
            seen["sweeps"] = max(seen["sweeps"], int(iters[k]))
            seen["worst_v"] = max(seen["worst_v"], rel(vn[6 * b0:6 * b1], v_ref))
            if c1 > c0:
                seen["worst_p"] = max(seen["worst_p"], rel(imp[c0:c1], imp_ref))
            seen["max_per_scene"] = max(seen["max_per_scene"], c1 - c0)
        seen["contacts"] += len(con["body_a"])
        # box face manifolds: 4 contact points of one body pair (or body and wall) sharing a normal
        key = np.concatenate([con["body_a"][:, None], con["body_b"][:, None],
                              np.round(con["frame"][:, :3] * 1e6).astype(np.int64)], axis=1)
        _, cnt = np.unique(key, axis=0, return_counts=True)
        seen["manifolds"] += int((cnt == 4).sum())

    device_steps(ctx, b, bd, q, v, prev, 5, dt, prm, check)
    print(f"launch {launch}: {seen['contacts']} contacts over 5 steps, max {seen['max_per_scene']}/scene, "
          f"{seen['manifolds']} 4-point manifolds, max sweeps {seen['sweeps']}, "
          f"worst rel v {seen['worst_v']:.2e} p {seen['worst_p']:.2e}, LDS {seen['lds']}")
    assert seen["worst_v"] <= TOL and seen["worst_p"] <= TOL
    assert seen["max_per_scene"] >= 100 and seen["manifolds"] > 0 and seen["sweeps"] > 10
    if launch[1] == 1:
        assert all(seen["lds"])  # the one-scene-per-workgroup launch ran the LDS kernel
