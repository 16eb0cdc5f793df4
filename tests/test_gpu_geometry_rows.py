"""GPU: the geometry form of the contact rows (cdx_set_contact_points).  k_geo_pack rebuilds every
rigid side's angular rows r_body x (R^T d_k) from the contact points and the bodies' positions and
quaternions and finds them equal, bit for bit, to the rows the host contact generator wrote (a
sphere pile with spun orientations, a sphere/box scene); the range sweep that rebuilds the rows in
every sweep then gives exactly the bits of the range sweep that reads them, after fixed sweeps and
to tolerance, row-scaled and block, and matches the oracle's pj_sweep (projected_jacobi.cpp:121-138)
to <= 1e-10; likewise the Chebyshev and Nesterov loops' range sweeps.  A row changed in its last bit is found
(the sweep then reads the rows), and new positions, new contacts or a Moreau drift discard the
points."""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import native, scenes
from test_geometry_rows import mixed_scene

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


@pytest.fixture
def gctx():
    c = native.Context(0)
    c.set_range_sweep(2)
    yield c
    c.close()


def pile_problem(n=3000, seed=4):
    pile = scenes.sphere_pile(n=n, settle_steps=1500, seed=seed)
    rng = np.random.default_rng(seed)
    q = pile.q.copy()
    qq = rng.normal(size=(len(q), 4))
    q[:, 3:] = qq / np.linalg.norm(qq, axis=1, keepdims=True)
    v = 0.2 * rng.normal(size=(len(q), 6))
    b = scenes.sphere_bodies(pile.radius, pile.mass, q, v.reshape(-1))
    return scenes.make_problem(b, pile.contacts(q=q))


def mixed_problem():
    q, dof, con = mixed_scene()
    b = scenes.batched_drop(nscenes=3, seed=1, nmin=20, nmax=40)
    rng = np.random.default_rng(9)
    bodies = scenes.shape_bodies(b.shape, b.dims, b.mass, q, 0.2 * rng.normal(size=(len(q), 6)).reshape(-1))
    return scenes.make_problem(bodies, con)


def upload(c, prob, points=True):
    c.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    c.set_contacts(*prob.contact_arrays())
    if points:
        c.set_contact_points(prob.point)


def solve(c, prob, points, **kw):
    upload(c, prob, points)
    info = c.pj_solve(5e-4, 1.0, params(**kw))
    return c.velocities(), c.impulses(), info, c.geometry_rows_info()


@pytest.mark.parametrize("accel", [native.CDX_ACCEL_CHEBYSHEV, native.CDX_ACCEL_NESTEROV], ids=["chebyshev", "nesterov"])
def test_accelerated_loops_same_bits(gctx, accel):
    """The Chebyshev and Nesterov loops' range sweeps with the rebuilt rows against the same loops
    reading them."""
    prob = pile_problem(n=2000, seed=6)
    for k, tol in ((1, 1e-15), (9, 1e-15), (20000, 1e-9)):
        kw = dict(max_iter=k, tol_abs=tol, tol_rel=tol, accel=accel)
        v1, i1, inf1, g1 = solve(gctx, prob, True, **kw)
        v0, i0, inf0, g0 = solve(gctx, prob, False, **kw)
        assert g1 == (True, 0, True) and g0[2] is False and gctx.range_sweep_info()[0]
        assert inf1.iters == inf0.iters
        np.testing.assert_array_equal(v1, v0)
        np.testing.assert_array_equal(i1, i0)


@pytest.mark.parametrize("build", [pile_problem, mixed_problem], ids=["pile", "spheres_boxes"])
@pytest.mark.parametrize("block", [0, 1])
def test_rebuilt_rows_give_the_same_bits(gctx, build, block):
    prob = build()
    assert prob.nc > 100
    upload(gctx, prob)
    assert gctx.geometry_rows_info()[:2] == (True, 0)
    for k, tol in ((1, 1e-15), (7, 1e-15), (40, 1e-15), (20000, 1e-9)):
        v1, i1, inf1, g1 = solve(gctx, prob, True, max_iter=k, tol_abs=tol, tol_rel=tol, block_precond=block)
        v0, i0, inf0, g0 = solve(gctx, prob, False, max_iter=k, tol_abs=tol, tol_rel=tol, block_precond=block)
        assert g1 == (True, 0, True) and g0[2] is False and gctx.range_sweep_info()[0]
        assert inf1.iters == inf0.iters
        np.testing.assert_array_equal(v1, v0)
        np.testing.assert_array_equal(i1, i0)
        if k <= 40:
            st, v, imp, res = oracle.pj_solve(prob, 5e-4, oracle.params(max_iter=k, tol_abs=tol, tol_rel=tol,
                                                                        block_precond=block))
            assert st == 0 and res.iters == inf1.iters == k
            assert rel(v1, v) <= TOL and rel(i1, imp) <= TOL


def test_a_changed_bit_is_found(gctx):
    prob = mixed_problem()
    rigid = np.flatnonzero((prob.body_a >= 0) & (prob.dof[np.maximum(prob.body_a, 0)] == 6))
    c = int(rigid[len(rigid) // 2])
    prob.angA = prob.angA.copy()
    prob.angA[c, 4] = np.nextafter(prob.angA[c, 4], np.inf)
    upload(gctx, prob)
    assert gctx.geometry_rows_info()[:2] == (False, 1)
    v1, i1, info, g = solve(gctx, prob, True, max_iter=9, tol_abs=1e-15, tol_rel=1e-15)
    assert g[2] is False  # the sweep read the rows
    st, v, imp, res = oracle.pj_solve(prob, 5e-4, oracle.params(max_iter=9, tol_abs=1e-15, tol_rel=1e-15))
    assert rel(v1, v) <= TOL and rel(i1, imp) <= TOL


def test_state_changes_discard_the_points(gctx):
    prob = mixed_problem()
    upload(gctx, prob)
    assert gctx.geometry_rows_info()[:2] == (True, 0)
    gctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    assert gctx.geometry_rows_info()[:2] == (False, -1)
    upload(gctx, prob)
    gctx.set_contacts(*prob.contact_arrays())
    assert gctx.geometry_rows_info()[:2] == (False, -1)
    gctx.set_contact_points(prob.point)
    gctx.moreau_predrift(5e-4, 0.5)  # positions move: the lever arms no longer hold
    assert gctx.geometry_rows_info()[:2] == (False, -1)
    gctx.set_contact_points(prob.point)  # the rows were written for the old positions
    assert gctx.geometry_rows_info()[0] is False
    upload(gctx, prob)
    gctx.set_geometry_rows(0)
    gctx.pj_solve(5e-4, 1.0, params(max_iter=3))
    assert gctx.geometry_rows_info() == (True, 0, False)
    with pytest.raises(native.CdxError):
        gctx.set_geometry_rows(2)
