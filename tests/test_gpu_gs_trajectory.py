"""Coloured Gauss-Seidel on the device against the oracle's sequential PGS, by trajectory.

Coloured GS visits the contact blocks colour by colour, the reference's PGS in DinvDiag order
(projected_gauss_seidel.cpp:144-169), so their iterates differ and are compared by what the
solution must satisfy instead: the kinetic-energy trajectory and the complementarity residual.
The workload is the 1000-sphere drop (r = 0.05 m, m = 1 kg, mu = 0.5, h = 1e-3 s, 200 Moreau
steps), both solvers run to tol 1e-8 every step, and each trajectory evolves from its own
solutions (contacts regenerated on the CPU from its own positions, warm start matched from its
own previous impulses).

Bars: the contact count of every step is the same on both trajectories; the kinetic energy agrees
to 1e-6 relative at every step (the oracle's own coloured sweep stays within 1.3e-9 of PGS here);
every device solve meets the natural-map complementarity residual |p - Pi(p - D^-1 gamma)| / |p|
<= 1e-4 (PGS reaches 2.1e-5 at worst); the sweeps summed over the trajectory within 10 % of PGS's
(single steps differ more: the sweep order changes which contact settles first).
"""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import scenes
from cardillocxx_amd.solver import Config, MoreauIntegrator

pytestmark = pytest.mark.gpu

STEPS, DT, TOL = 200, 1e-3, 1e-8
GS_TO_PJ = np.array([1.0, -1.0, -1.0])  # PGS stores tangents with the sign of its internal lambda


def kinetic_energy(minv, v):
    m = np.where(minv > 0, 1.0 / np.where(minv > 0, minv, 1.0), 0.0)
    return 0.5 * float((m * v * v).sum())


def pgs_trajectory(scene):
    b = scene.bodies()
    q, v, prev = b["q"].copy(), b["v"].copy(), None
    out = []
    for s in range(STEPS):
        con = scene.contacts(q.reshape(-1, 7))
        nc = len(con["body_a"])
        con["warm"] = scenes.match_warmstart(prev, con) if prev is not None and nc else np.zeros((nc, 3))
        prob = scenes.make_problem(dict(b, v=v, q=q), con)
        st, vn, imp, r = oracle.pgs_solve(prob, DT, oracle.params(tol_abs=TOL, tol_rel=TOL))
        assert st == 0
        comp = oracle.complementarity(prob, vn, imp * GS_TO_PJ) if nc else 0.0
        out.append((nc, kinetic_energy(b["minv"], vn), r.iters, comp))
        q = oracle.position_update(b["dof"], q, vn, DT)
        v = vn
        prev = dict(body_a=con["body_a"], body_b=con["body_b"], point=con["point"], imp=imp)
    return out


def test_colored_gs_kinetic_energy_trajectory_matches_pgs():
    scene = scenes.sphere_drop_grid()
    ref = pgs_trajectory(scene)
    b = scene.bodies()
    cfg = Config(friction_enable=True, pj_tol_abs=TOL, pj_tol_rel=TOL)
    integ = MoreauIntegrator(scene, cfg, solver="gs")
    worst_ke = worst_comp = 0.0
    max_contacts = sweeps = sweeps_ref = 0
    for s in range(STEPS):
        v_n = integ.ctx.velocities() if s > 0 else b["v"]
        q_n = integ.positions().reshape(-1)
        info = integ.step(DT)
        con = integ.last_contacts
        nc = len(con["body_a"])
        v = integ.ctx.velocities()
        nc_ref, ke_ref, it_ref, comp_ref = ref[s]
        assert nc == nc_ref, s
        ke = kinetic_energy(b["minv"], v)
        worst_ke = max(worst_ke, abs(ke - ke_ref) / ke_ref)
        if nc:
            prob = scenes.make_problem(dict(b, v=v_n, q=q_n), con)
            comp = oracle.complementarity(prob, v, con["imp"] * GS_TO_PJ)
            worst_comp = max(worst_comp, comp)
            assert comp <= 1e-4, (s, comp, comp_ref)
            sweeps += info.iters
            sweeps_ref += it_ref
        max_contacts = max(max_contacts, nc)
    print(f"coloured GS vs PGS over {STEPS} steps: max contacts {max_contacts}, worst rel KE {worst_ke:.2e}, "
          f"worst complementarity {worst_comp:.2e} (PGS {max(r[3] for r in ref):.2e}), sweeps {sweeps} (PGS {sweeps_ref})")
    assert max_contacts > 400
    assert worst_ke <= 1e-6
    assert abs(sweeps - sweeps_ref) <= 0.1 * sweeps_ref
