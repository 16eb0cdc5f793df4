"""CPU contact generation for spheres and boxes (the step before the device solve): closed-form
geometry of simple configurations, the contact conventions the solve relies on (normal from side A
to side B, orthonormal frames, body-frame angular rows), and bounded manifolds."""
import math

import numpy as np
import pytest

from cardillocxx_amd import scenes


def two_bodies(shape_a, shape_b, dims_a, dims_b, qa, qb):
    return scenes.ShapeBatch(body_off=np.array([0, 2], np.int32), shape=np.array([shape_a, shape_b], np.int32),
                             dims=np.array([dims_a, dims_b], float), mass=np.ones(2), mu_body=np.full(2, 0.5),
                             q=np.array([qa, qb], float), v=np.zeros((2, 6)),
                             walls=(np.zeros((0, 3)), np.zeros(0), np.zeros(0)), margin=1e-4)


def quat_z(angle):
    return [0.0, 0.0, math.sin(angle / 2), math.cos(angle / 2)]


def check_conventions(c, q):
    n = c["frame"][:, 0:3]
    for k in range(3):
        np.testing.assert_allclose(np.linalg.norm(c["frame"][:, 3 * k:3 * k + 3], axis=1), 1.0, atol=1e-12)
    # angular rows are r_body x d_body with r the arm from the body centre to the contact point
    for i in range(len(c["body_a"])):
        for side, ang in ((c["body_a"][i], c["angA"][i]), (c["body_b"][i], c["angB"][i])):
            if side < 0:
                continue
            qq = q[side]
            x, y, z, w = qq[3:]
            R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                          [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                          [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
            r = R.T @ (c["point"][i] - qq[:3])
            for k in range(3):
                np.testing.assert_allclose(ang[3 * k:3 * k + 3], np.cross(r, R.T @ c["frame"][i, 3 * k:3 * k + 3]),
                                           atol=1e-12)
    return n


def test_sphere_sphere_closed_form():
    b = two_bodies(scenes.SPHERE, scenes.SPHERE, [0.1, 0, 0], [0.2, 0, 0], [0, 0, 0, 0, 0, 0, 1],
                   [0.299, 0, 0, 0, 0, 0, 1])
    c = b.contacts()
    assert len(c["body_a"]) == 1 and c["body_a"][0] == 0 and c["body_b"][0] == 1
    n = check_conventions(c, b.q)
    np.testing.assert_allclose(n[0], [1, 0, 0], atol=1e-15)
    assert c["depth"][0] == pytest.approx(0.001, abs=1e-12)


def test_box_face_contact_on_a_rotated_box_keeps_four_points():
    """A box resting on a larger box, turned 45 degrees about the normal: the clipped face is an
    octagon (8 vertices); the manifold keeps 4, including the deepest, spanning the overlap."""
    h = 0.1
    tilt = 1e-3  # slight tilt so one corner is deepest
    qb = np.array([0, 0, 2 * h - 5e-5, 0, 0, 0, 1.0])
    qb[3:] = quat_z(math.pi / 4)
    rx = [math.sin(tilt / 2), 0, 0, math.cos(tilt / 2)]
    x1, y1, z1, w1 = qb[3:]
    x2, y2, z2, w2 = rx
    qb[3:] = [w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
              w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2]
    b = two_bodies(scenes.BOX, scenes.BOX, [0.12, 0.12, h], [h, h, h], [0, 0, 0, 0, 0, 0, 1], qb)
    c = b.contacts()
    m = len(c["body_a"])
    assert m == 4
    n = check_conventions(c, b.q)
    np.testing.assert_allclose(n, np.tile([0, 0, 1.0], (m, 1)), atol=1e-12)
    # the points span the overlap: their convex hull is far from degenerate
    p = c["point"][:, :2]
    span = max(np.linalg.norm(p[i] - p[j]) for i in range(m) for j in range(m))
    assert span > 0.15
    # the deepest vertex of the octagon is kept: the tilt about the top box's x axis sinks its
    # clipped vertices at body y = -0.1 by 0.1 sin(tilt) below the 5e-5 overlap at the centre
    assert c["depth"].max() == pytest.approx(5e-5 + 0.1 * math.sin(tilt), rel=1e-3)


def test_box_on_plane_four_corners():
    b = scenes.boxes_on_plane(nscenes=1, per_scene=1)
    c = b.contacts()
    assert len(c["body_a"]) == 4 and (c["body_a"] == -1).all() and (c["body_b"] == 0).all()
    check_conventions(c, b.q)


def test_batched_drop_manifolds_bounded():
    """No body pair of the batched drop carries more than four contacts."""
    b = scenes.batched_drop(nscenes=8, seed=3)
    rng = np.random.default_rng(0)
    q = b.q.copy()
    q[:, 2] = q[:, 2] * 0.3  # squeeze the scenes downwards so bodies overlap
    c = b.contacts(q=q)
    pairs = {}
    for a, bb in zip(c["body_a"], c["body_b"]):
        pairs[(a, bb)] = pairs.get((a, bb), 0) + 1
    box = lambda i: i >= 0 and b.shape[i] == scenes.BOX
    assert max(v for (a, bb), v in pairs.items() if box(a) and box(bb)) <= 4
    assert np.isfinite(c["frame"]).all() and np.isfinite(c["angA"]).all() and np.isfinite(c["angB"]).all()
