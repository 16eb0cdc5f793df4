"""Differentiable Moreau rollouts over the device solve.

A rollout of N steps maps (v0, per-body friction mu, per-body mass m) to the final body positions.
Each step is the reference Moreau step (moreau.cpp:6-13, theta = 1): CPU contacts at q_n, device
projected-Jacobi solve v_{n+1} = S(v_n; M^-1(m), f_ext(m), mu), q_{n+1} = q_n + h B(q_n) v_{n+1}.

Backward: the adjoint of each solve comes from the device (``cdx_pj_adjoint``: implicit
differentiation of the converged fixed point, or reverse mode through the recorded sweeps of a
cold-start re-solve to the same tolerance, the FP64 reference gradient), with the contact geometry
of every step held fixed (W_n, c_n from the forward trajectory).  Positions
obey x_{n+1} = x_n + h v_{n+1}, so dL/dx_N reaches every step's v_{n+1} as h dL/dx_N.

Memory: only (q, v) every ``checkpoint_every`` steps are kept from the forward pass.  The
backward re-runs each segment forward from its checkpoint (regenerating contacts), keeps that
segment's per-step state and converged impulses, and walks it in reverse; each step's device state
is restored by a one-sweep solve warm-started from its converged impulses.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import native
from .scenes import GRAVITY, ShapeBatch, shape_bodies


@dataclasses.dataclass
class Grads:
    v0: np.ndarray  # [nb, 6]
    mu: np.ndarray  # [nb]
    mass: np.ndarray  # [nb]


class Rollout:
    def __init__(self, batch: ShapeBatch, steps: int, dt: float = 1e-3, checkpoint_every: int = 10, tol: float = 1e-10,
                 adjoint_mode: int = native.CDX_ADJ_IMPLICIT, adjoint_tol: float = 1e-11, device: int = 0, ctx=None,
                 max_iter: int = 200000, batched: bool = False, adjoint_max_iter: int = 100000,
                 store_steps: bool = True):
        self.batch = batch
        self.steps = steps
        self.dt = dt
        self.k = max(1, checkpoint_every)
        self._own_ctx = ctx is None
        self.ctx = ctx or native.Context(device)
        self.prm = native.default_params()
        self.prm.tol_abs = self.prm.tol_rel = tol
        self.prm.max_iter = max_iter
        self.prm.warmstart = 1
        self.adjoint_mode = adjoint_mode
        self.adjoint_tol = adjoint_tol
        self.record_bytes = 32 << 30  # device memory for the sweeps recorded by the unrolled backward
        self.adjoint_max_iter = adjoint_max_iter
        # batched: every forward solve is one cdx_batch_pj_solve over the batch's scenes, each
        # scene converging on its own residual exactly as a standalone solve of that scene would
        # (the adjoints run on the restored state of each step, so every mode applies)
        self.batched = batched
        self.frozen = None  # optional per-step contact dicts to reuse (geometry held fixed)
        self.stats = dict(forward_solves=0, sweeps=0, adjoint_iters=0, unrolled_fallbacks=0)
        # store_steps: the forward keeps every step's record (contacts, q, v, impulses), so the
        # backward recomputes nothing; otherwise it recomputes each segment from its checkpoint
        self.store_steps = store_steps
        self.records = None

    def close(self):
        """Releases the device context if this rollout created it."""
        if self._own_ctx and self.ctx is not None:
            self.ctx.close()
            self.ctx = None

    # ---- model parameters -> solver inputs
    def _bodies(self, v, mass):
        b = self.batch
        bd = shape_bodies(b.shape, b.dims, mass, b.q, v)
        return bd

    def _contact_mu(self, con, mu_body, mu_wall):
        a, bb = con["body_a"], con["body_b"]
        ma = np.where(a >= 0, mu_body[np.maximum(a, 0)], np.inf)
        mb = np.where(bb >= 0, mu_body[np.maximum(bb, 0)], np.inf)
        mw = mu_wall if np.isscalar(mu_wall) else np.asarray(mu_wall).max()
        ma = np.where(a >= 0, ma, mw)
        mb = np.where(bb >= 0, mb, mw)
        mu = np.minimum(ma, mb)
        src = np.where(ma <= mb, a, bb)  # body whose friction is the min (-1: a wall)
        return mu, src

    def _contacts(self, step, q):
        if self.frozen is not None:
            return self.frozen[step]
        return self.batch.contacts(q=q.reshape(-1, 7))

    def _solve(self, q, v, minv, fext, fgyro, con, mu, warm=None, prm=None):
        ctx = self.ctx
        ctx.set_bodies(np.full(len(q) // 7, 6, np.int32), minv, v, fext, fgyro, q)
        ctx.set_contacts(con["body_a"], con["body_b"], mu, con["frame"], con["angA"], con["angB"], con["rest"],
                         con["vbias"], con["warm"] if warm is None else warm)
        self.stats["forward_solves"] += 1
        if self.batched and prm is None:
            ctx.set_scenes(self.batch.body_off, con["contact_off"])
            iters, _ = ctx.batch_pj_solve(self.dt, 1.0, self.prm)
            self.stats["sweeps"] += int(iters.max()) if len(iters) else 0
            return None
        info = ctx.pj_solve(self.dt, 1.0, prm or self.prm)
        self.stats["sweeps"] += info.iters
        return info

    def _step(self, step, q, v, minv, fext, mu_body, prev):
        """One forward step; returns (q_next, v_next, record)."""
        from .scenes import match_warmstart

        con = self._contacts(step, q)
        mu, src = self._contact_mu(con, mu_body, self.batch.walls[2])
        warm = match_warmstart(prev, con) if prev is not None and len(con["body_a"]) else np.zeros((len(con["body_a"]), 3))
        fgyro = _gyro(minv, v)
        self._solve(q, v, minv, fext, fgyro, con, mu, warm)
        v1 = self.ctx.velocities()
        imp = self.ctx.impulses()
        self.ctx.moreau_update(self.dt, 1.0)
        q1 = self.ctx.positions()
        rec = dict(q=q, v=v, con=con, mu=mu, src=src, warm=warm, imp=imp, fgyro=fgyro)
        nxt = dict(body_a=con["body_a"], body_b=con["body_b"], point=con["point"], imp=imp)
        return q1, v1, rec, nxt

    def forward(self, v0, mu_body, mass):
        """Runs the rollout; returns final positions [nb, 3]."""
        self.v0 = np.ascontiguousarray(v0, np.float64).reshape(-1)
        self.mu_body = np.ascontiguousarray(mu_body, np.float64)
        self.mass = np.ascontiguousarray(mass, np.float64)
        bd = self._bodies(self.v0, self.mass)
        self.minv, self.fext = bd["minv"], bd["fext"]
        q, v = bd["q"].copy(), self.v0.copy()
        self.checkpoints = {}
        self.contacts_log = []
        self.records = [] if self.store_steps else None
        prev = None
        for n in range(self.steps):
            if n % self.k == 0:
                self.checkpoints[n] = (q.copy(), v.copy(), prev)
            q, v, rec, prev = self._step(n, q, v, self.minv, self.fext, self.mu_body, prev)
            self.contacts_log.append(rec["con"])
            if self.records is not None:
                self.records.append(rec)
        self.qN = q
        return q.reshape(-1, 7)[:, :3].copy()

    def _step_adjoint(self, rec, lv):
        """Adjoint of one step's solve given dL/dv_{n+1} = lv.  Implicit: restore the converged
        state with one sweep from its impulses and iterate the adjoint.  AUTO: the same implicit
        system by restarted GMRES, which also converges where the contact impulses are not unique
        (a box face on four coplanar points) and returns the least-squares adjoint where the
        system is inconsistent (its relative residual is tracked in the stats).  Unrolled:
        reverse mode needs every sweep, so the step is re-solved from a cold start to the same
        tolerance with recording on.  Reverse mode holds the starting impulses constant, so a
        warm-started solve that converges within a few sweeps would drop how its warm start
        depends on the parameters; from p_0 = 0 the derivative of the sweep map converges to that
        of the fixed point."""
        mode = self.adjoint_mode
        if mode in (native.CDX_ADJ_IMPLICIT, native.CDX_ADJ_AUTO):
            prm = native.default_params()
            prm.max_iter, prm.tol_abs, prm.tol_rel, prm.warmstart = 1, 0.0, 0.0, 1
            self._solve(rec["q"], rec["v"], self.minv, self.fext, rec["fgyro"], rec["con"], rec["mu"], rec["imp"], prm)
            if self.batched:  # one Krylov space per scene in the GMRES adjoint
                self.ctx.set_scenes(self.batch.body_off, rec["con"]["contact_off"])
            g = self.ctx.pj_adjoint(self.dt, self.prm, lv, None, mode=mode, max_iter=self.adjoint_max_iter,
                                    tol=self.adjoint_tol)
            if mode == native.CDX_ADJ_AUTO and len(rec["con"]["body_a"]):
                r = float(g["info"].residual)
                self.stats["worst_adjoint_residual"] = max(self.stats.get("worst_adjoint_residual", 0.0), r)
                if r > self.adjoint_tol:
                    self.stats["least_squares_adjoints"] = self.stats.get("least_squares_adjoints", 0) + 1
            return g
        prm = native.default_params()
        prm.tol_abs, prm.tol_rel, prm.max_iter, prm.warmstart = self.prm.tol_abs, self.prm.tol_rel, self.prm.max_iter, 0
        per_sweep = 8 * (8 * len(self.mass) + 3 * len(rec["con"]["body_a"]))
        self.ctx.set_recording(True, max(1, min(prm.max_iter, self.record_bytes // per_sweep)))
        try:
            self._solve(rec["q"], rec["v"], self.minv, self.fext, rec["fgyro"], rec["con"], rec["mu"],
                        np.zeros_like(rec["warm"]), prm)
            return self.ctx.pj_adjoint(self.dt, self.prm, lv, None, mode=native.CDX_ADJ_UNROLLED)
        finally:
            self.ctx.set_recording(False)

    def backward(self, g_pos):
        """dL/d(v0, mu_body, mass) given dL/dx_N (positions, [nb, 3])."""
        nb = len(self.mass)
        gx = np.zeros((nb, 6))
        gx[:, :3] = np.asarray(g_pos, np.float64).reshape(nb, 3)
        gx = gx.reshape(-1)
        a = np.zeros(nb * 6)  # adjoint of v_{n+1} from later steps
        g_minv = np.zeros(nb * 6)
        g_fext = np.zeros(nb * 6)
        g_mu = np.zeros(nb)

        def segments():  # (first step, records of the segment's steps), last segment first
            if self.records is not None:
                yield 0, self.records
                return
            for s in reversed(sorted(self.checkpoints)):
                e = min(s + self.k, self.steps)
                q, v, prev = self.checkpoints[s]
                seg = []
                for n in range(s, e):  # recompute the segment forward from its checkpoint
                    q1, v1, rec, prev = self._step(n, q, v, self.minv, self.fext, self.mu_body, prev)
                    seg.append(rec)
                    q, v = q1, v1
                yield s, seg

        for s, seg in segments():
            for n in range(s + len(seg) - 1, s - 1, -1):
                rec = seg[n - s]
                g = self._step_adjoint(rec, a + self.dt * gx)
                self.stats["adjoint_iters"] += g["info"].iters
                a = g["v"]
                g_minv += g["minv"]
                g_fext += g["fext"]
                src = rec["src"]
                ok = src >= 0
                np.add.at(g_mu, src[ok], g["mu"][ok])
        # chain to the parameters: minv = [1/m]*3 + 1/(m i_unit), f_ext = m g
        minv = self.minv.reshape(nb, 6)
        g_mass = (g_minv.reshape(nb, 6) * (-minv / self.mass[:, None])).sum(axis=1)
        g_mass += (g_fext.reshape(nb, 6)[:, :3] * GRAVITY[None, :]).sum(axis=1)
        return Grads(v0=a.reshape(nb, 6), mu=g_mu, mass=g_mass)


@dataclasses.dataclass
class ChainGrads:
    v0: np.ndarray  # [nb, 6]
    damping: np.ndarray  # [nj] dL/dD_rot per joint
    mu: np.ndarray  # [nc_total] dL/dmu summed per contact of every step (diagnostic)


class ChainRollout(Rollout):
    """Rollout of bodies joined by damped spherical joints (joints.spherical_rows): joint rows and
    contacts in one projected Jacobi per step (cdx_pj_solve, or cdx_batch_pj_solve per scene),
    with Lambda carried from step to step.  Parameters: initial velocities v0 and the rotational
    damping D of every joint; masses and friction are fixed.  The backward uses the implicit
    adjoint through the joint rows (cdx_pj_adjoint_joints) with contact geometry and joint
    Jacobians of every step held fixed; a damper row has A = 1/D, so dL/dD = -dL/dA / D^2.
    As in Rollout, store_steps (default) keeps every step's state in the forward (the contacts and
    joint rows it keeps anyway, plus q, v and the impulses)."""

    def __init__(self, batch: ShapeBatch, joints_spec: dict, steps: int, **kw):
        super().__init__(batch, steps, **kw)
        from . import joints as jmod
        self.jmod = jmod
        self.spec = joints_spec
        self.frames = jmod.joint_frames(batch.q, joints_spec)
        self.nj = len(joints_spec["body_a"])
        self.frozen_rows = None  # optional per-step joint rows to reuse (geometry held fixed)

    def _rows(self, step, q):
        if self.frozen_rows is None:
            return self.jmod.spherical_rows(q, self.spec, self.frames, self.damping)
        r = dict(self.frozen_rows[step])  # frozen Jacobians; the damper rows follow the damping
        damper = np.arange(r["ns"] + r["nd"]) >= r["ns"]
        r["comp"] = r["comp"].copy()
        r["comp"][damper] = 1.0 / self.damping[r["joint"][damper]]
        return r

    def _solve_joints(self, q, v, fgyro, con, warm, rows, lam_prev, prm=None):
        ctx = self.ctx
        ctx.set_bodies(np.full(len(q) // 7, 6, np.int32), self.minv, v, self.fext, fgyro, q)
        ctx.set_contacts(con["body_a"], con["body_b"], con["mu"], con["frame"], con["angA"], con["angB"], con["rest"],
                         con["vbias"], warm)
        ctx.set_joints(rows, lam_prev=lam_prev)
        self.stats["forward_solves"] += 1
        if self.batched and prm is None:
            ctx.set_scenes(self.batch.body_off, con["contact_off"])
            iters, _ = ctx.batch_pj_solve(self.dt, 1.0, self.prm)
            self.stats["sweeps"] += int(iters.max()) if len(iters) else 0
            return
        info = ctx.pj_solve(self.dt, 1.0, prm or self.prm)
        self.stats["sweeps"] += info.iters

    def _chain_step(self, step, q, v, prev, lam_prev):
        from .scenes import match_warmstart

        con = self._contacts(step, q)
        warm = match_warmstart(prev, con) if prev is not None and len(con["body_a"]) else np.zeros((len(con["body_a"]), 3))
        rows = self._rows(step, q)
        fgyro = _gyro(self.minv, v)
        self._solve_joints(q, v, fgyro, con, warm, rows, lam_prev)
        v1, imp, lam = self.ctx.velocities(), self.ctx.impulses(), self.ctx.joint_impulses()
        self.ctx.moreau_update(self.dt, 1.0)
        q1 = self.ctx.positions()
        rec = dict(q=q, v=v, con=con, warm=warm, imp=imp, fgyro=fgyro, rows=rows, lam_prev=lam_prev)
        nxt = dict(body_a=con["body_a"], body_b=con["body_b"], point=con["point"], imp=imp)
        return q1, v1, rec, nxt, lam

    def forward(self, v0, damping):
        """Runs the rollout; returns final positions [nb, 3]."""
        b = self.batch
        self.v0 = np.ascontiguousarray(v0, np.float64).reshape(-1)
        self.damping = np.broadcast_to(np.asarray(damping, np.float64), (self.nj,)).copy()
        self.mass = np.asarray(b.mass, np.float64)
        bd = self._bodies(self.v0, self.mass)
        self.minv, self.fext = bd["minv"], bd["fext"]
        q, v = bd["q"].copy(), self.v0.copy()
        self.checkpoints = {}
        self.contacts_log, self.rows_log = [], []
        self.records = [] if self.store_steps else None
        prev, lam = None, None
        for n in range(self.steps):
            if n % self.k == 0:
                self.checkpoints[n] = (q.copy(), v.copy(), prev, lam)
            q, v, rec, prev, lam = self._chain_step(n, q, v, prev, lam)
            self.contacts_log.append(rec["con"])
            self.rows_log.append(rec["rows"])
            if self.records is not None:
                self.records.append(rec)
        self.qN = q
        return q.reshape(-1, 7)[:, :3].copy()

    def backward(self, g_pos):
        """dL/d(v0, damping) given dL/dx_N (positions, [nb, 3])."""
        nb = len(self.mass)
        gx = np.zeros((nb, 6))
        gx[:, :3] = np.asarray(g_pos, np.float64).reshape(nb, 3)
        gx = gx.reshape(-1)
        a = np.zeros(nb * 6)
        g_damp = np.zeros(self.nj)
        g_mu = []
        prm1 = native.default_params()
        prm1.max_iter, prm1.tol_abs, prm1.tol_rel, prm1.warmstart = 1, 0.0, 0.0, 1

        def segments():  # (first step, records of the segment's steps), last segment first
            if self.records is not None:
                yield 0, self.records
                return
            for s in reversed(sorted(self.checkpoints)):
                e = min(s + self.k, self.steps)
                q, v, prev, lam = self.checkpoints[s]
                seg = []
                for n in range(s, e):
                    q1, v1, rec, prev, lam = self._chain_step(n, q, v, prev, lam)
                    seg.append(rec)
                    q, v = q1, v1
                yield s, seg

        for s, seg in segments():
            for n in range(s + len(seg) - 1, s - 1, -1):
                rec = seg[n - s]
                # restore the converged state: one sweep warm-started from the converged impulses
                self._solve_joints(rec["q"], rec["v"], rec["fgyro"], rec["con"], rec["imp"], rec["rows"], rec["lam_prev"],
                                   prm1)
                g = self.ctx.pj_adjoint_joints(self.dt, self.prm, a + self.dt * gx, max_iter=self.adjoint_max_iter,
                                               tol=self.adjoint_tol, mode=native.CDX_ADJ_AUTO)
                self.stats["adjoint_iters"] += g["info"].iters
                if g["info"].residual > self.adjoint_tol:  # least-squares adjoint of an inconsistent system
                    self.stats["inexact_adjoints"] = self.stats.get("inexact_adjoints", 0) + 1
                    self.stats["worst_adjoint_residual"] = max(self.stats.get("worst_adjoint_residual", 0.0),
                                                               g["info"].residual)
                a = g["v"]
                rows = rec["rows"]
                damper = np.arange(rows["ns"] + rows["nd"]) >= rows["ns"]
                np.add.at(g_damp, rows["joint"][damper], -g["comp"][damper] / self.damping[rows["joint"][damper]] ** 2)
                g_mu.append(g["mu"])
        return ChainGrads(v0=a.reshape(nb, 6), damping=g_damp, mu=np.concatenate(g_mu[::-1]) if g_mu else np.zeros(0))


def _gyro(minv, v):
    nb = len(v) // 6
    w = v.reshape(nb, 6)[:, 3:]
    I = 1.0 / minv.reshape(nb, 6)[:, 3:]
    f = np.zeros((nb, 6))
    f[:, 3:] = -np.cross(w, I * w)
    return f.reshape(-1)


def com_loss(pos, mass, body_off, target):
    """sum_s || COM_s - target_s ||^2 and its gradients w.r.t. positions and masses."""
    ns = len(body_off) - 1
    L = 0.0
    gpos = np.zeros_like(pos)
    gm = np.zeros_like(mass)
    for s in range(ns):
        b0, b1 = body_off[s], body_off[s + 1]
        m = mass[b0:b1]
        M = m.sum()
        com = (m[:, None] * pos[b0:b1]).sum(0) / M
        r = com - target[s]
        L += float(r @ r)
        gpos[b0:b1] = 2.0 * r[None, :] * (m / M)[:, None]
        gm[b0:b1] = 2.0 * (pos[b0:b1] - com[None, :]) @ r / M
    return L, gpos, gm


try:  # the autograd wrapper needs torch; the rollout itself does not
    import torch

    class RolloutFunction(torch.autograd.Function):
        """torch.autograd.Function over a Rollout: final positions from (v0, mu, mass)."""

        @staticmethod
        def forward(ctx, v0, mu, mass, rollout):
            pos = rollout.forward(v0.detach().cpu().double().numpy(), mu.detach().cpu().double().numpy(),
                                  mass.detach().cpu().double().numpy())
            ctx.rollout = rollout
            return torch.from_numpy(pos)

        @staticmethod
        def backward(ctx, g_pos):
            g = ctx.rollout.backward(g_pos.detach().cpu().double().numpy())
            return torch.from_numpy(g.v0), torch.from_numpy(g.mu), torch.from_numpy(g.mass), None

    def rollout_loss(rollout, v0, mu, mass, target):
        """Differentiable loss sum_s ||COM_s(final) - target_s||^2 (torch tensors in, scalar out)."""
        pos = RolloutFunction.apply(v0, mu, mass, rollout)
        off = torch.as_tensor(rollout.batch.body_off, dtype=torch.long)
        seg = torch.repeat_interleave(torch.arange(len(off) - 1), off[1:] - off[:-1])
        msum = torch.zeros(len(off) - 1, dtype=mass.dtype).index_add_(0, seg, mass)
        mx = torch.zeros(len(off) - 1, 3, dtype=pos.dtype).index_add_(0, seg, mass[:, None] * pos)
        com = mx / msum[:, None]
        return ((com - target) ** 2).sum()
except ImportError:  # pragma: no cover
    pass
