"""Step-level backward with checkpointing: gradients of sum_s ||COM_s(final) - target_s||^2
w.r.t. initial velocities, friction coefficients and masses, against central finite differences
of the same rollout with the contact geometry of every step held at the nominal trajectory's, and
implicit vs unrolled (FP64) per-step adjoints."""
import numpy as np
import pytest

from cardillocxx_amd import diff, native, scenes

pytestmark = pytest.mark.gpu


def setup(steps=12, k=5, mode=native.CDX_ADJ_IMPLICIT):
    batch = scenes.boxes_on_plane(nscenes=2, per_scene=4, seed=3, vstd=0.5)
    r = diff.Rollout(batch, steps=steps, dt=1e-3, checkpoint_every=k, tol=1e-12, adjoint_mode=mode, adjoint_tol=1e-13)
    rng = np.random.default_rng(0)
    target = rng.normal(scale=0.1, size=(batch.nscenes, 3))
    return batch, r, target


def loss_of(r, batch, target, v0, mu, mass):
    pos = r.forward(v0, mu, mass)
    return diff.com_loss(pos, mass, batch.body_off, target)


@pytest.mark.parametrize("scene,mode", [("boxes", native.CDX_ADJ_IMPLICIT), ("spheres", native.CDX_ADJ_IMPLICIT),
                                        ("spheres", native.CDX_ADJ_UNROLLED)])
def test_rollout_gradients_vs_finite_differences(scene, mode):
    if scene == "boxes":
        batch, r, target = setup(mode=mode)
    else:
        batch = spheres_on_plane()
        target = np.random.default_rng(0).normal(scale=0.1, size=(batch.nscenes, 3))
        r = diff.Rollout(batch, steps=12, dt=1e-3, checkpoint_every=5, tol=1e-12, adjoint_mode=mode, adjoint_tol=1e-13)
    v0, mu, mass = batch.v.reshape(-1).copy(), batch.mu_body.copy(), batch.mass.copy()
    L, gpos, gm_direct = loss_of(r, batch, target, v0, mu, mass)
    g = r.backward(gpos)
    r.frozen = r.contacts_log  # FD of the same map: geometry of every step held fixed
    rng = np.random.default_rng(1)
    for name, base, grad in (("v0", v0, g.v0.reshape(-1)), ("mu", mu, g.mu), ("mass", mass, g.mass + gm_direct)):
        for _ in range(2):
            d = rng.normal(size=base.shape)
            if name == "v0":
                d.reshape(-1, 6)[:, 2] = 0.0  # no vertical perturbation keeps the bodies on the plane
                if scene == "boxes":
                    d.reshape(-1, 6)[:, 3:] = 0.0
            eps = 1e-6

            def f(x):
                args = dict(v0=v0, mu=mu, mass=mass)
                args[name] = x
                return loss_of(r, batch, target, args["v0"], args["mu"], args["mass"])[0]

            fd = (f(base + eps * d) - f(base - eps * d)) / (2 * eps)
            an = float(grad @ d)
            assert an == pytest.approx(fd, rel=1e-4, abs=1e-9), (scene, name, an, fd)


def spheres_on_plane(nscenes=2, per_scene=4, seed=3, radius=0.05):
    """The box scene with every box replaced by a sphere: one contact per body, so the converged
    impulses are unique (a box face on 4 corners is statically indeterminate: its normal impulse
    split depends on the warm start, and the derivative of that split is path dependent)."""
    b = scenes.boxes_on_plane(nscenes=nscenes, per_scene=per_scene, seed=seed, vstd=0.5)
    b.shape[:] = scenes.SPHERE
    b.dims[:] = 0.0
    b.dims[:, 0] = radius
    b.q[:, 2] = radius
    b.v[:, 3:] = np.random.default_rng(seed + 1).normal(scale=2.0, size=(len(b.mass), 3))
    return b


def test_implicit_vs_unrolled_rollout_gradient():
    batch = spheres_on_plane()
    rng = np.random.default_rng(0)
    target = rng.normal(scale=0.1, size=(batch.nscenes, 3))
    kw = dict(steps=8, dt=1e-3, checkpoint_every=4, tol=1e-12, adjoint_tol=1e-13)
    ri = diff.Rollout(batch, adjoint_mode=native.CDX_ADJ_IMPLICIT, **kw)
    ru = diff.Rollout(batch, adjoint_mode=native.CDX_ADJ_UNROLLED, **kw)
    assert len(batch.contacts()["body_a"]) == batch.mass.size  # one sphere-plane contact each
    v0, mu, mass = batch.v.reshape(-1).copy(), batch.mu_body.copy(), batch.mass.copy()
    _, gpos, _ = loss_of(ri, batch, target, v0, mu, mass)
    gi = ri.backward(gpos)
    loss_of(ru, batch, target, v0, mu, mass)
    gu = ru.backward(gpos)
    for a, b in ((gi.v0, gu.v0), (gi.mu, gu.mu), (gi.mass, gu.mass)):
        np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-9 * max(1.0, np.abs(a).max()))


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    batch, r, target = setup(steps=6, k=3)
    v0 = torch.tensor(batch.v.reshape(-1, 6), dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(batch.mu_body, dtype=torch.float64, requires_grad=True)
    mass = torch.tensor(batch.mass, dtype=torch.float64, requires_grad=True)
    L = diff.rollout_loss(r, v0, mu, mass, torch.tensor(target))
    L.backward()
    L2, gpos, gm_direct = loss_of(r, batch, target, batch.v.reshape(-1), batch.mu_body, batch.mass)
    g = r.backward(gpos)
    assert float(L) == pytest.approx(L2, rel=1e-12)
    np.testing.assert_allclose(v0.grad.numpy(), g.v0, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(mu.grad.numpy(), g.mu, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(mass.grad.numpy(), g.mass + gm_direct, rtol=1e-10, atol=1e-14)


def test_batched_rollout_matches_scene_by_scene():
    """One batched solve per step (every scene on its own residual) gives each scene the
    trajectory and gradients of that scene rolled out alone through the grid-wide solve."""
    batch = spheres_on_plane(nscenes=3, per_scene=4)
    rng = np.random.default_rng(2)
    target = rng.normal(scale=0.1, size=(batch.nscenes, 3))
    kw = dict(steps=6, dt=1e-3, checkpoint_every=3, tol=1e-10, adjoint_tol=1e-13)
    rb = diff.Rollout(batch, batched=True, **kw)
    v0, mu, mass = batch.v.reshape(-1).copy(), batch.mu_body.copy(), batch.mass.copy()
    pos = rb.forward(v0, mu, mass)
    _, gpos, _ = diff.com_loss(pos, mass, batch.body_off, target)
    g = rb.backward(gpos)
    for s in range(batch.nscenes):
        b0, b1 = batch.body_off[s], batch.body_off[s + 1]
        sub = batch.scene(s)
        r1 = diff.Rollout(sub, batched=True, **kw)
        p1 = r1.forward(sub.v.reshape(-1).copy(), sub.mu_body.copy(), sub.mass.copy())
        np.testing.assert_allclose(p1, pos[b0:b1], rtol=0, atol=1e-13)
        g1 = r1.backward(gpos[b0:b1])
        np.testing.assert_allclose(g1.v0, g.v0[b0:b1], rtol=1e-8, atol=1e-12)
        np.testing.assert_allclose(g1.mu, g.mu[b0:b1], rtol=1e-8, atol=1e-12)
        np.testing.assert_allclose(g1.mass, g.mass[b0:b1], rtol=1e-8, atol=1e-12)


def test_auto_gmres_equals_implicit_iteration_where_impulses_are_unique():
    """AUTO solves the implicit adjoint system by restarted GMRES; where the fixed-point iteration
    converges (spheres: unique impulses) both give the same gradients."""
    batch = spheres_on_plane()
    target = np.random.default_rng(0).normal(scale=0.1, size=(batch.nscenes, 3))
    kw = dict(steps=6, dt=1e-3, checkpoint_every=3, tol=1e-12, adjoint_tol=1e-13)
    ra = diff.Rollout(batch, adjoint_mode=native.CDX_ADJ_AUTO, **kw)
    ri = diff.Rollout(batch, adjoint_mode=native.CDX_ADJ_IMPLICIT, **kw)
    v0, mu, mass = batch.v.reshape(-1).copy(), batch.mu_body.copy(), batch.mass.copy()
    _, gpos, _ = loss_of(ra, batch, target, v0, mu, mass)
    ga = ra.backward(gpos)
    loss_of(ri, batch, target, v0, mu, mass)
    gi = ri.backward(gpos)
    assert ra.stats["worst_adjoint_residual"] <= 1e-13
    for a, b in ((ga.v0, gi.v0), (ga.mu, gi.mu), (ga.mass, gi.mass)):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12 * max(1.0, np.abs(b).max()))


def test_auto_boxes_64_per_scene_vs_finite_differences():
    """The differentiable workload's scenes at reduced count: 4 scenes of 64 boxes on a plane
    (faces on four coplanar contacts: impulses not unique, where the fixed-point adjoint iteration
    drifts), 20 steps; AUTO (GMRES) gradients against central finite differences of the rollout
    with the geometry of every step held fixed, and the batched rollout against scene-by-scene
    rollouts through the same per-scene solve.  (Non-unique impulses make each trajectory depend
    on the rounding of its solves -- a box's friction limits follow its share of the normal
    impulse -- so scene-by-scene rollouts through the grid-wide solve, which sums in another
    order, follow slightly different trajectories with their own, equally valid, gradients.)"""
    batch = scenes.boxes_on_plane(nscenes=4, per_scene=64, seed=0)
    target = np.random.default_rng(0).normal(scale=0.1, size=(batch.nscenes, 3)) + np.array([0.875, 0.875, 0.05])
    kw = dict(steps=20, dt=1e-3, checkpoint_every=10, tol=1e-12, adjoint_mode=native.CDX_ADJ_AUTO, adjoint_tol=1e-12)
    r = diff.Rollout(batch, **kw)
    v0, mu, mass = batch.v.reshape(-1).copy(), batch.mu_body.copy(), batch.mass.copy()
    L, gpos, gm_direct = loss_of(r, batch, target, v0, mu, mass)
    g = r.backward(gpos)
    print(f"AUTO adjoint on 4 x 64 boxes: worst relative residual {r.stats.get('worst_adjoint_residual', 0):.2e}, "
          f"{r.stats.get('least_squares_adjoints', 0)} least-squares steps, {r.stats['adjoint_iters']} applications")
    r.frozen = r.contacts_log
    rng = np.random.default_rng(1)
    for name, base, grad in (("v0", v0, g.v0.reshape(-1)), ("mu", mu, g.mu), ("mass", mass, g.mass + gm_direct)):
        for _ in range(2):
            d = rng.normal(size=base.shape)
            if name == "v0":
                d.reshape(-1, 6)[:, 2] = 0.0
                d.reshape(-1, 6)[:, 3:] = 0.0
            eps = 1e-6

            def f(x):
                args = dict(v0=v0, mu=mu, mass=mass)
                args[name] = x
                return loss_of(r, batch, target, args["v0"], args["mu"], args["mass"])[0]

            fd = (f(base + eps * d) - f(base - eps * d)) / (2 * eps)
            an = float(grad @ d)
            assert an == pytest.approx(fd, rel=1e-4, abs=1e-9), (name, an, fd)
    r.frozen = None
    r.close()
    rb = diff.Rollout(batch, batched=True, **kw)
    loss_of(rb, batch, target, v0, mu, mass)
    gb = rb.backward(gpos)
    worst = 0.0
    for s in (0, 3):
        b0, b1 = batch.body_off[s], batch.body_off[s + 1]
        sub = batch.scene(s)
        r1 = diff.Rollout(sub, batched=True, **kw)
        _, gp1, _ = loss_of(r1, sub, target[s][None, :], sub.v.reshape(-1).copy(), sub.mu_body.copy(), sub.mass.copy())
        g1 = r1.backward(gp1)
        # the masses' gradient through the solve is ~1e-18 here (equal masses on a plane): it is
        # measured against the scene's other gradients
        scale = max(np.abs(gb.v0[b0:b1]).max(), np.abs(gb.mu[b0:b1]).max())
        for a, c in ((g1.v0, gb.v0[b0:b1]), (g1.mu, gb.mu[b0:b1]), (g1.mass, gb.mass[b0:b1])):
            worst = max(worst, float(np.abs(a - c).max() / max(np.abs(c).max(), 1e-3 * scale)))
    print(f"batched vs scene-by-scene gradients: max relative difference {worst:.2e}")
    assert worst <= 1e-8


@pytest.mark.parametrize("mode", [native.CDX_ADJ_IMPLICIT, native.CDX_ADJ_AUTO])
def test_stored_steps_equal_checkpoint_recompute(mode):
    """Backward from the forward's stored step records = the checkpointed backward that recomputes
    each segment (the recomputed steps are the same solves), bitwise."""
    grads = []
    for store in (True, False):
        batch = scenes.boxes_on_plane(nscenes=2, per_scene=4, seed=3, vstd=0.5)
        r = diff.Rollout(batch, steps=12, dt=1e-3, checkpoint_every=5, tol=1e-12, adjoint_mode=mode, adjoint_tol=1e-13,
                         store_steps=store)
        target = np.random.default_rng(0).normal(scale=0.1, size=(batch.nscenes, 3))
        _, gpos, _ = loss_of(r, batch, target, batch.v.reshape(-1).copy(), batch.mu_body.copy(), batch.mass.copy())
        grads.append(r.backward(gpos))
        assert (r.records is not None) == store
    for key in ("v0", "mu", "mass"):
        np.testing.assert_array_equal(getattr(grads[0], key), getattr(grads[1], key))
