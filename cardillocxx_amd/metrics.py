"""Solution-quality metrics reported next to throughput.

complementarity(): natural-map residual of the frictional contact problem with diagonal Delassus
scaling, ||p - Pi_K(p - D^-1 gamma(v))|| / ||p|| over all contact rows, where gamma is the
post-solve contact velocity W v + c + e (W v_n + c) (the same bias as pj_assembler.cpp:127-170)
and Pi_K projects normals to >= 0 and tangents into the Coulomb disk mu p_n
(projected_jacobi.cpp:100-119).  Row conventions as dynamics_assembler.cpp:236-312: side A enters
with -1, side B with +1, each row [s d ; s r x d]; static sides contribute nothing.  Impulses use
the projected-Jacobi tangent sign (tangential components as the Jacobi solver stores them).
Vectorised over contacts of 6-dof bodies.
"""
from __future__ import annotations

import numpy as np


def _side_rows(prob, v, body, ang, sign):
    """Per contact and row k: s * (d_k . v_lin + ang_k . v_ang) and s^2 * (d_k^2 m^-1 + ang_k^2 I^-1)."""
    nc = len(body)
    d = prob.frame.reshape(nc, 3, 3)
    a = ang.reshape(nc, 3, 3)
    ok = body >= 0
    bi = np.where(ok, body, 0)
    vb = v.reshape(-1, 6)[bi]
    mb = prob.minv.reshape(-1, 6)[bi]
    g = sign * (np.einsum("ckj,cj->ck", d, vb[:, :3]) + np.einsum("ckj,cj->ck", a, vb[:, 3:]))
    D = np.einsum("ckj,cj->ck", d * d, mb[:, :3]) + np.einsum("ckj,cj->ck", a * a, mb[:, 3:])
    return np.where(ok[:, None], g, 0.0), np.where(ok[:, None], D, 0.0)


def complementarity(prob, v, imp) -> float:
    if not (np.asarray(prob.dof) == 6).all():
        raise ValueError("complementarity: 6-dof bodies only")
    a = np.asarray(prob.body_a)
    b = np.asarray(prob.body_b)
    live = (a >= 0) | (b >= 0)
    gA, DA = _side_rows(prob, v, a, prob.angA, -1.0)
    gB, DB = _side_rows(prob, v, b, prob.angB, 1.0)
    g0A, _ = _side_rows(prob, prob.v, a, prob.angA, -1.0)
    g0B, _ = _side_rows(prob, prob.v, b, prob.angB, 1.0)
    c = prob.vbias.reshape(-1, 3)
    e = prob.rest.reshape(-1, 3)
    gam = gA + gB + c + e * (g0A + g0B + c)
    D = DA + DB
    p = np.array(imp, dtype=np.float64).reshape(-1, 3)
    fr = prob.mu > 0
    p[~fr, 1:] = 0.0
    z = p - np.where(D > 0, gam / np.where(D > 0, D, 1.0), 0.0)
    z[:, 0] = np.maximum(z[:, 0], 0.0)
    tn = np.sqrt(z[:, 1] ** 2 + z[:, 2] ** 2)
    s = np.where(tn > 0, np.minimum(1.0, prob.mu * z[:, 0] / np.where(tn > 0, tn, 1.0)), 1.0)
    z[:, 1:] *= s[:, None]
    z[~fr, 1:] = 0.0
    r = np.where(live[:, None], p - z, 0.0)
    den = np.sqrt((np.where(live[:, None], p, 0.0) ** 2).sum())
    return float(np.sqrt((r ** 2).sum()) / max(den, 1e-300))
