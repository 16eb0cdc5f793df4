"""Bilateral joint rows (damped spherical joints of capsule chains) and the oracle's
extended-system projected Jacobi: the rows against the time derivative of the joint offset, the
oracle's solve against every block equation of S x = rhs, and a chain resting on a heightfield at
convergence (contact complementarity, joint velocity constraints, momentum balance)."""
import math

import numpy as np
import pytest

import oracle
from cardillocxx_amd import joints, metrics, scenes


def chain(nchains=2, links=5, seed=0, n_terrain=33):
    t = scenes.fractal_heightfield(n=n_terrain, extent=1.6, amp=0.03, seed=seed)
    batch, jt = scenes.capsule_chains(nchains=nchains, links=links, terrain=t, seed=seed)
    return batch, jt


def random_state(batch, seed):
    rng = np.random.default_rng(seed)
    q = batch.q.copy()
    qq = q[:, 3:] + 0.3 * rng.normal(size=(len(q), 4))  # bend the chain
    q[:, 3:] = qq / np.linalg.norm(qq, axis=1, keepdims=True)
    v = rng.normal(size=(len(q), 6))
    return q, v


def test_spring_rows_are_the_time_derivative_of_the_joint_offset():
    batch, jt = chain()
    q, v = random_state(batch, 1)
    frames = joints.joint_frames(batch.q, jt)
    rows = joints.spherical_rows(q, jt, frames, damping=0.0)
    assert rows["nd"] == 0 and rows["ns"] == 3 * len(jt["body_a"])
    dof = np.full(len(q), 6, np.int32)
    h = 1e-6
    qp = oracle.position_update(dof, q.reshape(-1), v.reshape(-1), h).reshape(-1, 7)
    qm = oracle.position_update(dof, q.reshape(-1), v.reshape(-1), -h).reshape(-1, 7)
    gp = joints.spherical_rows(qp, jt, frames, 0.0)["gerr"]
    gm = joints.spherical_rows(qm, jt, frames, 0.0)["gerr"]
    a, b = rows["body_a"], rows["body_b"]
    Gv = (rows["wa"] * v[a]).sum(1) + (rows["wb"] * v[b]).sum(1)
    np.testing.assert_allclose(Gv, (gp - gm) / (2 * h), rtol=1e-6, atol=1e-8 * np.abs(Gv).max())


def test_damper_rows_measure_relative_rotation():
    batch, jt = chain()
    q, _ = random_state(batch, 2)
    frames = joints.joint_frames(batch.q, jt)
    rows = joints.spherical_rows(q, jt, frames, damping=0.7)
    ns, nd = rows["ns"], rows["nd"]
    assert nd == ns and np.allclose(rows["comp"][ns:], 1 / 0.7) and (rows["comp"][:ns] == 0).all()
    R = joints.quat_to_rot(q[:, 3:])
    Om = np.array([0.3, -1.2, 0.8])
    v = np.zeros((len(q), 6))
    v[:, 3:] = np.einsum("nji,j->ni", R, Om)  # one world angular velocity, in each body's frame
    a, b = rows["body_a"][ns:], rows["body_b"][ns:]
    Gv = (rows["wa"][ns:] * v[a]).sum(1) + (rows["wb"][ns:] * v[b]).sum(1)
    np.testing.assert_allclose(Gv, 0.0, atol=1e-14)


def extended_system(prob, rows, dt, theta=1.0, lam_prev=None, beta=0.001):
    """S and rhs of PjAssembler::buildAndFactorS / rhs, assembled densely in numpy."""
    nv = int(np.sum(prob.dof))
    ns, nd = rows["ns"], rows["nd"]
    n = ns + nd
    lam_prev = np.zeros(n) if lam_prev is None else lam_prev
    G = np.zeros((n, nv))
    for r in range(n):
        G[r, 6 * rows["body_a"][r]:6 * rows["body_a"][r] + 6] += rows["wa"][r]
        G[r, 6 * rows["body_b"][r]:6 * rows["body_b"][r] + 6] += rows["wb"][r]
    M = 1.0 / prob.minv
    E = np.concatenate([rows["comp"][:ns] / (theta * dt * dt), rows["comp"][ns:] / (theta * dt)])
    S = np.block([[np.diag(M), G.T], [G, -np.diag(E)]])
    rhs = np.concatenate([M * prob.v + dt * prob.fext + dt * prob.fgyro, np.zeros(n)])
    Gv = G @ prob.v
    rhs[nv:nv + ns] = -(rows["comp"][:ns] * lam_prev[:ns]) / (theta * dt * dt) - (1 - theta) / theta * Gv[:ns] \
        - rows["vsrc"][:ns] / theta - (-(rows["comp"][:ns] * lam_prev[:ns]) / dt + rows["gerr"][:ns]) * (beta / (dt * theta))
    rhs[nv + ns:] = -(1 - theta) / theta * Gv[ns:] - rows["vsrc"][ns:] / theta
    return S, rhs, G


def chain_problem(batch, jt, q, v, damping, with_contacts):
    bd = scenes.shape_bodies(batch.shape, batch.dims, batch.mass, q, v.reshape(-1))
    con = batch.contacts(q=q) if with_contacts else scenes.shape_contacts(batch.body_off, q, batch.shape, batch.dims,
                                                                          batch.mu_body, batch.walls)
    prob = scenes.make_problem(bd, con)
    rows = joints.spherical_rows(q, jt, joints.joint_frames(batch.q, jt), damping)
    return prob, rows, con


@pytest.mark.parametrize("theta", [1.0, 0.6])
def test_oracle_extended_solve_satisfies_S(theta):
    batch, jt = chain()
    q, v = random_state(batch, 3)
    q[:, 2] += 1.0  # far above the terrain: no contacts
    prob, rows, con = chain_problem(batch, jt, q, 0.1 * v, damping=0.4, with_contacts=True)
    assert prob.nc == 0
    lam_prev = np.random.default_rng(4).normal(size=rows["ns"] + rows["nd"])
    st, vo, imp, lam, res = oracle.pj_solve_joints(prob, rows, 1e-3, oracle.params(), theta=theta, lam_prev=lam_prev)
    assert st == 0
    S, rhs, G = extended_system(prob, rows, 1e-3, theta=theta, lam_prev=lam_prev)
    x = np.concatenate([vo, lam])
    np.testing.assert_allclose(S @ x, rhs, rtol=0, atol=1e-9 * np.abs(rhs).max())
    np.testing.assert_allclose(x, np.linalg.solve(S, rhs), rtol=1e-8, atol=1e-12 * np.abs(x).max())
    # hard joints (C = 0), theta = 1: the joint velocity is the Baumgarte target -beta g / dt
    if theta == 1.0:
        ns = rows["ns"]
        np.testing.assert_allclose(G[:ns] @ vo, -0.001 * rows["gerr"][:ns] / 1e-3, atol=1e-12)


def test_oracle_chain_on_terrain_converged():
    batch, jt = chain(nchains=2, links=6, seed=5)
    q = batch.q.copy()
    q[:, 2] = scenes.terrain_height(batch.terrain, q[:, 0], q[:, 1]) + 0.02 - 2e-4
    v = np.zeros((len(q), 6))
    v[:, 0] = 0.2
    prob, rows, con = chain_problem(batch, jt, q, v, damping=0.05, with_contacts=True)
    assert prob.nc >= len(q)
    dt = 1e-3
    st, vo, imp, lam, res = oracle.pj_solve_joints(prob, rows, dt, oracle.params(tol_abs=1e-12, tol_rel=1e-12))
    assert st == 0 and res.residual <= 1
    assert metrics.complementarity(prob, vo, imp) < 1e-6
    S, rhs, G = extended_system(prob, rows, dt)
    ns = rows["ns"]
    np.testing.assert_allclose(G[:ns] @ vo, -0.001 * rows["gerr"][:ns] / dt, atol=1e-9)
    # momentum: M (v - v_n) - h f = W^T p - G^T Lambda  (top block of S with the contact impulses)
    W = np.zeros((3 * prob.nc, len(vo)))
    for c in range(prob.nc):
        for k in range(3):
            if prob.body_a[c] >= 0:
                W[3 * c + k, 6 * prob.body_a[c]:6 * prob.body_a[c] + 6] -= np.r_[prob.frame[c, 3 * k:3 * k + 3], prob.angA[c, 3 * k:3 * k + 3]]
            if prob.body_b[c] >= 0:
                W[3 * c + k, 6 * prob.body_b[c]:6 * prob.body_b[c] + 6] += np.r_[prob.frame[c, 3 * k:3 * k + 3], prob.angB[c, 3 * k:3 * k + 3]]
    top = (vo - prob.v) / prob.minv - dt * prob.fext - dt * prob.fgyro - W.T @ imp.reshape(-1) + G.T @ lam
    np.testing.assert_allclose(top, 0.0, atol=1e-10)
