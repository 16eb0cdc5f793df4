"""Bilateral joint rows for the extended projected-Jacobi system.

A damped spherical joint is the reference's TranslationRotationConstraint with infinite
translational stiffness and zero rotational stiffness (constraints.cpp:107-160):

* 3 spring rows (the translational part of the joint Jacobian) with compliance C = 1/K = 0, kept;
* 3 rotational spring rows with K = 0, i.e. C = inf, dropped (rows with C >= 1/EPS_C are not
  emitted, dynamics_assembler.cpp:407-418);
* damper rows "mirror g rows" (:143-145) with A = 1/D: the rotational ones for D_rot > 0 are
  kept, the translational ones (D_trans = 0, A = inf) dropped (:420-431).

Rows follow DynamicsAssembler::rebuildInteractionW_ (dynamics_assembler.cpp:341-439): all spring
rows of all joints (joint order), then all damper rows; per row the body-A and body-B
coefficients on [v; omega] are columns of WgA / WgB (buildJointJacobian, :107-127):

    translational row i:  A: [-A_IJ e_i ; (-[K1_r]x A_K1J - A_K1J [g]x) e_i]
                          B: [ A_IJ e_i ; ([K2_r]x A_K2J) e_i]
    rotational row i:     A: [0 ; A_K1J e_i]      B: [0 ; -A_K2J e_i]

with A_IJ = R_A A_K1J, A_K2J = R_B^T A_IJ and the joint offset g = A_IJ^T (x_J2 - x_J1)
(JointProperties::compute_g), which is also the position error of the translational rows
(getPositionError, :162-164).  Angular velocities are in body frames, as everywhere on this path.
"""
from __future__ import annotations

import numpy as np

EPS_C = 1e-10  # dynamics_assembler.cpp:364-365


def quat_to_rot(q):
    """Rotation matrices [n, 3, 3] of quaternions (x, y, z, w) [n, 4]."""
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((len(q), 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def skew(v):
    """[n, 3] -> [n, 3, 3] with skew(v) u = v x u (SkewSymmetricMatrix3r)."""
    S = np.zeros((len(v), 3, 3))
    S[:, 0, 1], S[:, 0, 2] = -v[:, 2], v[:, 1]
    S[:, 1, 0], S[:, 1, 2] = v[:, 2], -v[:, 0]
    S[:, 2, 0], S[:, 2, 1] = -v[:, 1], v[:, 0]
    return S


def joint_frames(q, joints):
    """A_K1J of each joint: the world-aligned joint frame (JointFrame default, A_refJ = I)
    expressed in body A at creation, A_K1J = R_A^T (transform::rotation, constraints.hpp:79-87)."""
    qq = np.asarray(q, dtype=np.float64).reshape(-1, 7)
    return np.transpose(quat_to_rot(qq[joints["body_a"], 3:]), (0, 2, 1)).copy()


def spherical_rows(q, joints, frames, damping):
    """Rows of damped spherical joints at positions q ([nb, 7]).

    joints: body_a, body_b (ints), anchor_a, anchor_b ([nj, 3], joint point in each body frame);
    frames: A_K1J [nj, 3, 3] (fixed at creation, see joint_frames);
    damping: rotational damping D_rot per joint (scalar or [nj]); 0 disables the damper rows.
    Returns a dict in the layout of cdx_set_joints: ns, nd, body_a, body_b, wa, wb ([n, 6]),
    comp, vsrc, gerr (springs), plus `joint` (the joint index of every row)."""
    qq = np.asarray(q, dtype=np.float64).reshape(-1, 7)
    a = np.asarray(joints["body_a"])
    b = np.asarray(joints["body_b"])
    nj = len(a)
    D = np.broadcast_to(np.asarray(damping, dtype=np.float64), (nj,))
    RA, RB = quat_to_rot(qq[a, 3:]), quat_to_rot(qq[b, 3:])
    K1 = np.asarray(joints["anchor_a"], dtype=np.float64)
    K2 = np.asarray(joints["anchor_b"], dtype=np.float64)
    AK1J = np.asarray(frames, dtype=np.float64)
    AIJ = RA @ AK1J
    AK2J = np.transpose(RB, (0, 2, 1)) @ AIJ
    xJ1 = qq[a, :3] + np.einsum("nij,nj->ni", RA, K1)
    xJ2 = qq[b, :3] + np.einsum("nij,nj->ni", RB, K2)
    g = np.einsum("nji,nj->ni", AIJ, xJ2 - xJ1)
    angA = -skew(K1) @ AK1J - AK1J @ skew(g)
    angB = skew(K2) @ AK2J
    # spring rows: 3 per joint, joint-major
    wa_s = np.concatenate([-AIJ, angA], axis=1).transpose(0, 2, 1).reshape(-1, 6)  # row i = column i
    wb_s = np.concatenate([AIJ, angB], axis=1).transpose(0, 2, 1).reshape(-1, 6)
    keep = D > 0
    A = np.where(keep, 1.0 / np.where(keep, D, 1.0), np.inf)
    zero = np.zeros((nj, 3, 3))
    wa_d = np.concatenate([zero, AK1J], axis=1).transpose(0, 2, 1)[keep].reshape(-1, 6)
    wb_d = np.concatenate([zero, -AK2J], axis=1).transpose(0, 2, 1)[keep].reshape(-1, 6)
    ns, nd = 3 * nj, 3 * int(keep.sum())
    assert np.all(A[keep] < 1.0 / EPS_C)
    jid = np.concatenate([np.repeat(np.arange(nj), 3), np.repeat(np.arange(nj)[keep], 3)])
    return dict(ns=ns, nd=nd,
                body_a=np.concatenate([np.repeat(a, 3), np.repeat(a[keep], 3)]).astype(np.int32),
                body_b=np.concatenate([np.repeat(b, 3), np.repeat(b[keep], 3)]).astype(np.int32),
                wa=np.ascontiguousarray(np.concatenate([wa_s, wa_d])), wb=np.ascontiguousarray(np.concatenate([wb_s, wb_d])),
                comp=np.concatenate([np.zeros(ns), np.repeat(A[keep], 3)]),
                vsrc=np.zeros(ns + nd), gerr=np.concatenate([g.reshape(-1), np.zeros(nd)]), joint=jid)
