"""Pins the oracle's Nesterov loop (nesterov_loop, projected_jacobi.cpp:379-449) independently.

Three checks, each written from the mathematics rather than from the oracle's code:

* a scalar FISTA recurrence (Beck & Teboulle 2009: t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2,
  y_{k+1} = z_{k+1} + (t_k - 1) / t_{k+1} (z_{k+1} - z_k)) on one frictionless point-mass contact,
  where a projected-Jacobi sweep scales the normal velocity by g = 1 - alpha while the impulse
  stays positive, so the k-th iterate has a closed form in g and the t-sequence;
* a dense restatement (numpy matrices, not the oracle's sparse rows) of the loop with its four
  restart rules -- residual growth by more than 5 %, beta above pj.nesterov_beta_threshold,
  the gradient-direction test (y - x_{k+1}) . (x_{k+1} - x_k) > 0 of O'Donoghue & Candes (2015),
  and momentum switched off after pj.nesterov_restart_limit restarts -- plus the non-finite
  residual rule, on problems chosen so that every rule fires; the oracle must follow its
  trajectory sweep by sweep and stop at the same iteration;
* mutation tests: copies of the oracle built with one planted mistake each (theta or beta
  recurrence, growth factor 1.5 instead of 1.05, reversed direction test, restart limit off by
  one, extrapolation from the wrong anchor) must fail the two checks above.
"""
import math

import numpy as np
import pytest

import oracle
from cardillocxx_amd import scenes
from test_oracle import G, delassus, two_planes

DT = 1e-3


def dense_nesterov(prob, dt, alpha=0.3, tol=1e-10, thr=1.0, limit=10000, max_iter=400):
    """Nesterov-accelerated projected Jacobi on dense matrices.  Returns (iterations, x, impulses
    per contact [nc, 3], set of restart rules that fired)."""
    W, c, order = delassus(prob)
    minv = prob.minv
    rows = [(ci, k) for ci in order for k in range(3 if prob.mu[ci] > 0 else 1)]
    e = np.array([prob.rest[ci, k] for ci, k in rows])
    D = (W ** 2) @ minv
    R = np.where(D > 0, alpha / np.where(D > 0, D, 1.0), 0.0)
    b = R * (c + e * (W @ prob.v + c))
    p0 = np.array([prob.warm[ci, k] for ci, k in rows])
    x0 = prob.v + dt * minv * (prob.fext + prob.fgyro) + minv * (W.T @ p0)

    def project(pt):
        out = pt.copy()
        i = 0
        while i < len(rows):
            mu = prob.mu[rows[i][0]]
            out[i] = max(pt[i], 0.0)
            if mu > 0:
                t = pt[i + 1:i + 3]
                tn = math.sqrt(t @ t)
                if tn > 0:
                    out[i + 1:i + 3] = t * min(1.0, mu * out[i] / tn)
                i += 3
            else:
                i += 1
        return out

    def sweep(x, p):
        pn = project(p - (R * (W @ x) + b))
        return x + minv * (W.T @ (pn - p)), pn

    def residual(a, bb):
        q = tol * np.maximum(np.abs(a), np.abs(bb)) + tol
        with np.errstate(divide="ignore", invalid="ignore"):
            r = (a - bb) / q
            return math.sqrt(r @ r) / math.sqrt(len(a))

    xk, pk, xy, py = x0, p0, x0, p0
    t, err_prev, restarts, off, fired = 1.0, math.inf, 0, False, set()
    it = 0
    for it in range(max_iter):
        x1, p1 = sweep(xy, py)
        err = residual(x1, xy)
        if err <= 1:
            break
        t1 = 0.5 * (1 + math.sqrt(1 + 4 * t * t))
        beta = (t - 1) / t1
        rule = None
        if not math.isfinite(err):
            rule = "nonfinite"
        elif err_prev < math.inf and err > 1.05 * err_prev:
            rule = "growth"
        elif beta > thr:
            rule = "beta"
        elif (xy - x1) @ (x1 - xk) > 0:
            rule = "direction"
        if rule:
            fired.add(rule)
            xy, py, xk, pk, t = x1, p1, x1, p1, 1.0
            restarts += 1
            if restarts >= limit:
                if not off:
                    fired.add("limit")
                off = True
        elif off:
            xy, py, xk, pk, t = x1, p1, x1, p1, 1.0
        else:
            xy, py, xk, pk, t = x1 + beta * (x1 - xk), p1 + beta * (p1 - pk), x1, p1, t1
        err_prev = err
    imp = np.zeros((prob.nc, 3))
    for (ci, k), val in zip(rows, p1):
        imp[ci, k] = val
    return it + 1, x1, imp, fired


def point_mass_contact(mu=0.0, m=1.5, v=(0.4, 0.1, -1.0)):
    """A point mass pressed into a tilted static plane (frame n, t1, t2 orthonormal, so W W^T = I)."""
    n = np.array([0.3, -0.2, 1.0])
    n /= np.linalg.norm(n)
    t1 = np.cross(n, [1.0, 0, 0])
    t1 /= np.linalg.norm(t1)
    return scenes.ContactProblem(
        dof=np.array([3], np.int32), minv=np.full(3, 1 / m), v=np.array(v, float), fext=np.array([0, 0, -m * G]),
        fgyro=np.zeros(3), q=np.zeros(3), body_a=np.array([-1], np.int32), body_b=np.array([0], np.int32),
        mu=np.array([mu]), frame=np.array([np.r_[n, t1, np.cross(n, t1)]]), angA=np.zeros((1, 9)),
        angB=np.zeros((1, 9)), rest=np.zeros((1, 3)), vbias=np.zeros((1, 3)), warm=np.zeros((1, 3)))


# (name, problem builder, dense/oracle options)
CASES = [
    ("two_planes_60_direction", lambda: two_planes(angle=60.0)[0], dict(alpha=0.3)),
    ("two_planes_30_growth", lambda: two_planes(angle=30.0)[0], dict(alpha=0.6)),
    ("two_planes_30_growth_a09", lambda: two_planes(angle=30.0)[0], dict(alpha=0.9)),
    ("stick_beta_threshold", lambda: point_mass_contact(mu=2.0), dict(alpha=0.3, thr=0.5)),
    ("stick_restart_limit", lambda: point_mass_contact(mu=2.0), dict(alpha=0.3, limit=2)),
    ("slip_direction", lambda: point_mass_contact(mu=0.1), dict(alpha=0.3)),
    ("zero_tolerance_nonfinite", lambda: point_mass_contact(mu=0.5), dict(alpha=0.3, tol=0.0, max_iter=12)),
] + [(f"random_{s}", (lambda s=s: scenes.random_problem(nb=4, nc=8, seed=s, p_point=0.5)),
      dict(alpha=0.3, max_iter=3000)) for s in range(6)]


def oracle_run(prob, opts, k, L=None):
    return oracle.pj_solve(prob, DT, oracle.params(alpha=opts.get("alpha", 0.3), max_iter=k, tol_abs=opts.get("tol", 1e-10),
                                                   tol_rel=opts.get("tol", 1e-10), nesterov=1,
                                                   nesterov_beta_threshold=opts.get("thr", 1.0),
                                                   nesterov_restart_limit=opts.get("limit", 10000)), L=L)


def rel(a, b, scale=0.0):
    """max |a - b| relative to max |b| (or to `scale` when larger: a converged sticking contact
    leaves velocities at rounding level, so those are measured against the free velocities)."""
    return np.abs(a - b).max() / max(np.abs(b).max(), scale, 1e-300)


def scalar_failures(L=None):
    """One frictionless contact, no restart in the first 5 iterations: the normal velocity after k
    sweeps is z_k with z_{k+1} = g y_k, y_0 = z_0 = eps0, y_k = z_k + beta_k (z_k - z_{k-1})."""
    prob = point_mass_contact(mu=0.0)
    m, alpha = 1.5, 0.3
    n = prob.frame[0, :3]
    eps0 = n @ (prob.v + DT * prob.fext / m)
    g = 1 - alpha
    fails = []
    z_prev, y, t = eps0, eps0, 1.0
    for k in range(1, 6):
        z = g * y
        st, v, imp, res = oracle_run(prob, dict(alpha=alpha, tol=1e-14), k, L)
        if not (st == 0 and res.iters == k and abs(imp[0, 0] - m * (z - eps0)) <= 1e-13 * abs(m * eps0)
                and abs(n @ v - z) <= 1e-13 * abs(eps0)):
            fails.append(f"scalar k={k}")
        t1 = 0.5 * (1 + math.sqrt(1 + 4 * t * t))
        y, z_prev, t = z + (t - 1) / t1 * (z - z_prev), z, t1
    return fails


def dense_failures(L=None, fired_out=None):
    fails = []
    for name, build, opts in CASES:
        prob = build()
        kw = {k: v for k, v in opts.items()}
        iters, x, imp, fired = dense_nesterov(prob, DT, **kw)
        if fired_out is not None:
            fired_out[name] = fired
        st, v, imp_o, res = oracle_run(prob, opts, kw.get("max_iter", 400), L)
        vs = np.abs(prob.v).max()
        if st != 0 or res.iters != iters or rel(v, x, vs) > 1e-12 or rel(imp_o, imp) > 1e-12:
            fails.append(f"{name} to tolerance: iters {res.iters} vs {iters}")
            continue
        for k in sorted({1, 2, 3, 5, 8, 13, 21, 34, iters - 1}):
            if not 1 <= k < iters:
                continue
            kk = dict(kw, max_iter=k)
            _, xk, impk, _ = dense_nesterov(prob, DT, **kk)
            st, v, imp_o, res = oracle_run(prob, dict(opts, max_iter=k), k, L)
            if st != 0 or res.iters != k or rel(v, xk, vs) > 1e-12 or rel(imp_o, impk) > 1e-12:
                fails.append(f"{name} after {k} sweeps")
                break
    return fails


def test_scalar_fista_recurrence():
    assert scalar_failures() == []


def test_dense_restatement_every_restart_rule():
    fired = {}
    assert dense_failures(fired_out=fired) == []
    seen = set().union(*fired.values())
    assert {"growth", "beta", "direction", "limit", "nonfinite"} <= seen, seen
    assert "beta" in fired["stick_beta_threshold"] and "limit" in fired["stick_restart_limit"]
    assert fired["two_planes_30_growth"] == {"growth"}


@pytest.mark.parametrize("mut", ["NES_THETA", "NES_BETA", "NES_GROWTH", "NES_DIRECTION", "NES_LIMIT", "NES_ANCHOR"])
def test_planted_mistakes_fail_the_pins(mut):
    L = oracle.mutant(mut)
    assert scalar_failures(L) + dense_failures(L), f"mutant {mut} passes every pin"
