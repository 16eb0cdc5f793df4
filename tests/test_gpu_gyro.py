"""Device gyroscopic force f_gyro = -omega x (I omega) (world.cpp:152-173) on non-isotropic bodies.

cdx_set_bodies with fgyro = NULL computes f_gyro on the device (k_gyro), and cdx_moreau_update
recomputes it from the new velocities.  A contact-free solve returns
x = v + h M^-1 (f_ext + f_gyro) (projected_jacobi.cpp:460-467, pj_assembler.cpp:150-153), so with
a large step and fast spins the solved velocities expose f_gyro: it moves them by more than 1 %,
and a wrong sign, axis or inertia index would fail the 1e-13 bar.  The oracle's gyro_force is
pinned by its closed-form test in test_oracle.py.
"""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import native, scenes

pytestmark = pytest.mark.gpu


def rel(a, b):
    den = max(np.abs(b).max(initial=0.0), 1e-300)
    return np.abs(np.asarray(a) - np.asarray(b)).max(initial=0.0) / den


def spinning_bodies(seed):
    """Unequal-edge boxes, capsules and spheres spinning at up to 20 rad/s about random axes,
    plus point masses (3 dof, no angular rows)."""
    rng = np.random.default_rng(seed)
    n = 96
    shape = np.array([scenes.BOX, scenes.CAPSULE, scenes.SPHERE] * (n // 3), np.int32)
    dims = np.zeros((n, 3))
    box = shape == scenes.BOX
    cap = shape == scenes.CAPSULE
    sph = shape == scenes.SPHERE
    dims[box] = rng.uniform(0.02, 0.15, (box.sum(), 3))  # half extents, all different
    dims[cap, 0] = rng.uniform(0.01, 0.04, cap.sum())     # radius
    dims[cap, 1] = rng.uniform(0.03, 0.12, cap.sum())     # half length
    dims[sph, 0] = rng.uniform(0.02, 0.08, sph.sum())
    mass = rng.uniform(0.2, 3.0, n)
    q = np.zeros((n, 7))
    q[:, :3] = rng.uniform(0, 1, (n, 3))
    qq = rng.normal(size=(n, 4))
    q[:, 3:] = qq / np.linalg.norm(qq, axis=1, keepdims=True)
    v = np.concatenate([rng.normal(0, 0.5, (n, 3)), rng.uniform(-20, 20, (n, 3))], axis=1)
    bd = scenes.shape_bodies(shape, dims, mass, q, v)
    # a few point masses appended (3 dof, 3 position entries)
    npnt = 8
    dof = np.concatenate([bd["dof"], np.full(npnt, 3, np.int32)])
    minv = np.concatenate([bd["minv"], np.full(3 * npnt, 2.0)])
    vv = np.concatenate([bd["v"], rng.normal(0, 1, 3 * npnt)])
    fext = np.concatenate([bd["fext"], np.tile([0.0, 0.0, -9.81 * 0.5], npnt)])
    qp = np.concatenate([bd["q"], rng.uniform(0, 1, 3 * npnt)])
    return dof, minv, vv, fext, qp


def free_solve(ctx, dt):
    p = native.default_params()
    info = ctx.pj_solve(dt, 1.0, p)
    assert info.iters == 0  # no contacts: the free solve, no sweeps
    return ctx.velocities()


def oracle_free(dof, minv, v, fext, dt, with_gyro=True):
    fg = oracle.gyro_force(dof, minv, v) if with_gyro else np.zeros_like(v)
    prob = scenes.ContactProblem(dof=dof, minv=minv, v=v, fext=fext, fgyro=fg, q=np.zeros(1),
                                 body_a=np.zeros(0, np.int32), body_b=np.zeros(0, np.int32), mu=np.zeros(0),
                                 frame=np.zeros((0, 9)), angA=np.zeros((0, 9)), angB=np.zeros((0, 9)),
                                 rest=np.zeros((0, 3)), vbias=np.zeros((0, 3)), warm=np.zeros((0, 3)))
    st, v_ref, _, res = oracle.pj_solve(prob, dt, oracle.params())
    assert st == 0
    return v_ref


@pytest.mark.parametrize("seed", [0, 1])
def test_device_gyro_force_in_free_solve_and_after_update(ctx, seed):
    dof, minv, v, fext, q = spinning_bodies(seed)
    dt = 0.02
    empty = np.zeros(0, np.int32)
    ctx.set_bodies(dof, minv, v, fext, None, q)
    ctx.set_contacts(empty, empty, np.zeros(0), np.zeros((0, 9)), np.zeros((0, 9)), np.zeros((0, 9)),
                     np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3)))
    v_n = v
    for step in range(3):
        v_dev = free_solve(ctx, dt)
        v_ref = oracle_free(dof, minv, v_n, fext, dt)
        v_nogyro = oracle_free(dof, minv, v_n, fext, dt, with_gyro=False)
        # f_gyro matters here: leaving it out moves the solved velocities by more than 1 %
        assert rel(v_nogyro, v_ref) > 1e-2
        assert rel(v_dev, v_ref) <= 1e-13, step
        ctx.moreau_update(dt, 1.0)  # v_n <- v_dev, q advanced, f_gyro recomputed on the device
        v_n = v_dev
