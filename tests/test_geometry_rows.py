"""The host contact generator (libcdx_scene) writes a rigid side's angular rows r_body x (R^T d_k),
r_body = R^T (p - x_com) (makeContact + buildContactRowByDof, contact_jacobian.hpp:17-23), with plain
IEEE operations in a fixed order and no contraction, so that the device can rebuild them bit for
bit from the contact points (cdx_set_contact_points, the geometry form of the range sweep).  Here,
on the CPU: the rows of a sphere pile and of a mixed sphere/box scene with random orientations equal
a numpy evaluation of the same operations (numpy ufuncs round every operation and never contract),
the second tangent is n x t1 bit for bit, and a planted change of the operation order breaks that
equality.  The device side of the claim
(k_geo_pack finds no mismatching contact) is in tests/test_gpu_geometry_rows.py."""
import numpy as np
import pytest

from cardillocxx_amd import scenes


def rot_t(q, v, swap=False):
    """R^T v for unit quaternions q = (qx, qy, qz, qw) [n, 4], entries of R first."""
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    r00, r01, r02 = 1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)
    r10, r11, r12 = 2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)
    r20, r21, r22 = 2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)
    if swap:  # the same sum in another order: a plausible way to lose bit equality
        return np.stack([r00 * v[:, 0] + (r10 * v[:, 1] + r20 * v[:, 2]), r01 * v[:, 0] + r11 * v[:, 1] + r21 * v[:, 2],
                         r02 * v[:, 0] + r12 * v[:, 1] + r22 * v[:, 2]], 1)
    return np.stack([r00 * v[:, 0] + r10 * v[:, 1] + r20 * v[:, 2], r01 * v[:, 0] + r11 * v[:, 1] + r21 * v[:, 2],
                     r02 * v[:, 0] + r12 * v[:, 1] + r22 * v[:, 2]], 1)


def cross(a, b):
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1)


def rebuilt_rows(q7, dof, con, swap=False):
    """(side, contact index, rebuilt 9 angular entries) for every rigid side."""
    out = []
    for side, key in ((0, "body_a"), (1, "body_b")):
        ids = con[key]
        sel = np.flatnonzero(ids >= 0)
        sel = sel[dof[ids[sel]] == 6]
        qb = q7[ids[sel]]
        r = rot_t(qb[:, 3:], con["point"][sel] - qb[:, :3], swap)
        ang = np.concatenate([cross(r, rot_t(qb[:, 3:], con["frame"][sel, 3 * k:3 * k + 3], swap)) for k in range(3)], 1)
        out.append((side, sel, ang))
    return out


def mixed_scene():
    b = scenes.batched_drop(nscenes=3, seed=1, nmin=20, nmax=40)
    q = b.q.copy()
    q[:, 2] = 0.03 + 0.3 * q[:, 2]  # pressed together and onto the floor: box-box, box-wall, sphere contacts
    q[:, :2] = 0.5 + 0.6 * (q[:, :2] - 0.5)
    con = scenes.shape_contacts(b.body_off, q, b.shape, b.dims, b.mu_body, b.walls, b.margin, True)
    dof = scenes.shape_bodies(b.shape, b.dims, b.mass, q)["dof"]
    return q, np.asarray(dof), con


def pile_scene():
    pile = scenes.sphere_pile(n=400, settle_steps=1000, seed=2)
    rng = np.random.default_rng(3)
    q = pile.q.copy()
    qq = rng.normal(size=(len(q), 4))  # spun: the lever arms are not along the world axes
    q[:, 3:] = qq / np.linalg.norm(qq, axis=1, keepdims=True)
    con = pile.contacts(q=q)
    return q, np.full(len(q), 6), con


@pytest.mark.parametrize("build", [mixed_scene, pile_scene], ids=["spheres_boxes", "sphere_pile"])
def test_rows_are_the_plain_ieee_evaluation(build):
    q, dof, con = build()
    assert len(con["body_a"]) > 50
    f = con["frame"]
    # the normal row is the unit normal the frame was built from, so t2 = n x t1 exactly
    np.testing.assert_array_equal(cross(f[:, 0:3], f[:, 3:6]).view(np.int64), f[:, 6:9].view(np.int64))
    np.testing.assert_allclose(np.linalg.norm(f[:, 0:3], axis=1), 1.0, rtol=0, atol=4e-16)
    checked = 0
    for side, sel, ang in rebuilt_rows(q, dof, con):
        given = (con["angA"] if side == 0 else con["angB"])[sel]
        np.testing.assert_array_equal(ang.view(np.int64), given.view(np.int64))  # bit for bit, signed zeros too
        checked += len(sel)
    assert checked > 50


def test_another_operation_order_is_caught():
    q, dof, con = mixed_scene()
    differ = 0
    for side, sel, ang in rebuilt_rows(q, dof, con, swap=True):
        given = (con["angA"] if side == 0 else con["angB"])[sel]
        differ += int((ang.view(np.int64) != given.view(np.int64)).any(axis=1).sum())
    assert differ > 0
