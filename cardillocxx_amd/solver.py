"""Host-side mirror of the reference solver interface, over the C ABI.

* :class:`Config` -- the hot-path keys of cardillo::config::Config (src/config/config.hpp:46-113,232).
* :class:`SolverBase` -- cardillo::solver::SolverBase (src/physics/solver/solver_base.hpp:14-36):
  ``solve(dt, theta)`` returns the stacked velocities; ``last_iterations()``, ``last_error()``,
  ``name()``.  Failures raise (the reference throws ``std::runtime_error``).
* :class:`ProjectedJacobiSolver`, :class:`ColoredGaussSeidelSolver` -- the device backends that drop
  in for ProjectedJacobiSolver (projected_jacobi.hpp:120-163) and ProjectedGaussSeidelSolver.
* :class:`MoreauIntegrator` -- MoreauIntegrator::step (src/physics/integration/moreau.cpp:6-13)
  around a scene whose contacts are regenerated on the CPU every step.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import native
from .scenes import ContactProblem, match_warmstart


@dataclasses.dataclass
class Config:
    friction_enable: bool = False  # friction.enable (config.hpp:46)
    pj_max_iterations: int = 200000  # pj.max_iterations
    pj_tol_abs: float = 1e-5
    pj_tol_rel: float = 1e-5
    pj_relaxation: float = 1.0
    pj_alpha: float = 0.3
    pj_nesterov: bool = False
    pj_nesterov_beta_threshold: float = 1.0
    pj_nesterov_restart_limit: int = 10000
    pj_chebyshev: bool = False  # pj.chebyshev
    pj_anderson: bool = False  # pj.anderson
    pj_anderson_m: int = 2  # pj.anderson_m
    pj_rdiag_true_delassus: bool = False  # pj.rdiag_true_delassus
    pj_warmstart: bool = True
    moreau_theta: float = 1.0
    moreau_implicit_gyroscopy: bool = False  # moreau.implicit_gyroscopy (config.hpp:233)
    collision_match_max_dist: float = 0.02
    pj_check_every: int = 10  # residual cadence of the PJ standard loop (projected_jacobi.cpp:375)

    def params(self) -> native.SolverParams:
        p = native.default_params()
        p.alpha = self.pj_alpha
        p.relaxation = self.pj_relaxation
        p.max_iter = max(1, self.pj_max_iterations)
        p.tol_abs = self.pj_tol_abs
        p.tol_rel = self.pj_tol_rel
        p.check_every = self.pj_check_every
        p.warmstart = 1 if self.pj_warmstart else 0
        # precedence pj_nesterov > pj_chebyshev > pj_anderson (projected_jacobi.cpp:469-472)
        p.accel = (native.CDX_ACCEL_NESTEROV if self.pj_nesterov else native.CDX_ACCEL_CHEBYSHEV if self.pj_chebyshev
                   else native.CDX_ACCEL_ANDERSON if self.pj_anderson else native.CDX_ACCEL_NONE)
        p.anderson_m = max(1, self.pj_anderson_m)
        p.block_precond = 1 if self.pj_rdiag_true_delassus else 0
        p.nesterov_beta_threshold = self.pj_nesterov_beta_threshold
        p.nesterov_restart_limit = self.pj_nesterov_restart_limit
        p.implicit_gyro = 1 if self.moreau_implicit_gyroscopy else 0
        return p


class SolverBase:
    """Mirror of cardillo::solver::SolverBase."""

    def __init__(self):
        self._last_iters = -1
        self._last_error = 0.0

    def solve(self, dt: float, theta: float) -> np.ndarray:  # pragma: no cover - abstract
        raise NotImplementedError

    def last_iterations(self) -> int:
        return self._last_iters

    def last_error(self) -> float:
        return self._last_error

    def name(self) -> str:
        return "SolverBase"


class _DeviceSolver(SolverBase):
    def __init__(self, problem: ContactProblem, cfg: Config, ctx: native.Context | None = None):
        super().__init__()
        self.problem = problem
        self.cfg = cfg
        self.ctx = ctx or native.Context(0)
        self.info = None

    def _upload(self):
        p = self.problem
        mu = p.mu if self.cfg.friction_enable else np.zeros_like(p.mu)
        self.ctx.set_bodies(p.dof, p.minv, p.v, p.fext, p.fgyro, p.q)
        self.ctx.set_contacts(p.body_a, p.body_b, mu, p.frame, p.angA, p.angB, p.rest, p.vbias, p.warm)

    def _finish(self, info):
        self.info = info
        self._last_iters = info.iters
        self._last_error = info.residual
        # storeImpulse (warmstart.hpp:41-56): the impulses become the contacts' last_impulse
        self.problem.warm = self.ctx.impulses()
        return self.ctx.velocities()


class ProjectedJacobiSolver(_DeviceSolver):
    def solve(self, dt: float, theta: float) -> np.ndarray:
        self._upload()
        return self._finish(self.ctx.pj_solve(dt, theta, self.cfg.params()))

    def name(self) -> str:
        return "ProjectedJacobi"


class ColoredGaussSeidelSolver(_DeviceSolver):
    def solve(self, dt: float, theta: float) -> np.ndarray:
        self._upload()
        p = self.cfg.params()
        # ProjectedGaussSeidelSolver reads pj.nesterov only (projected_gauss_seidel.cpp:196)
        p.accel = native.CDX_ACCEL_NESTEROV if self.cfg.pj_nesterov else native.CDX_ACCEL_NONE
        return self._finish(self.ctx.gs_colored_solve(dt, theta, p))

    def name(self) -> str:
        return "ColoredGaussSeidel"


class MoreauIntegrator:
    """Moreau theta-method over a SphereScene: CPU contacts at q_{n+theta}, device solve, device
    velocity write-back and position drift (moreau.cpp:6-13)."""

    def __init__(self, scene, cfg: Config, solver: str = "pj", device: int = 0):
        self.scene = scene
        self.cfg = cfg
        self.solver = solver
        self.ctx = native.Context(device)
        b = scene.bodies()
        self.dof = b["dof"]
        self.ctx.set_bodies(b["dof"], b["minv"], b["v"], b["fext"], None, b["q"])
        self.prev = None
        self.last_info = None
        self.last_problem = None

    def positions(self):
        return self.ctx.positions().reshape(-1, 7)

    def step(self, dt: float):
        theta = self.cfg.moreau_theta
        self.ctx.moreau_predrift(dt, theta)
        q = self.positions()
        con = self.scene.contacts(q, friction=self.cfg.friction_enable)
        if self.prev is not None and self.cfg.pj_warmstart:
            con["warm"] = match_warmstart(self.prev, con, self.cfg.collision_match_max_dist)
        self.ctx.set_contacts(con["body_a"], con["body_b"], con["mu"], con["frame"], con["angA"], con["angB"],
                              con["rest"], con["vbias"], con["warm"])
        prm = self.cfg.params()
        if self.solver == "pj":
            info = self.ctx.pj_solve(dt, theta, prm)
        else:
            prm.accel = native.CDX_ACCEL_NESTEROV if self.cfg.pj_nesterov else native.CDX_ACCEL_NONE
            info = self.ctx.gs_colored_solve(dt, theta, prm)
        con["imp"] = self.ctx.impulses()
        self.prev = con
        self.last_info = info
        self.last_contacts = con
        self.ctx.moreau_update(dt, theta)
        return info
