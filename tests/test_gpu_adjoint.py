"""Gradients of the device PJ solve: implicit adjoint and unrolled reverse mode against central
finite differences of the device forward (contact geometry held fixed, FP64)."""
import numpy as np
import pytest

from cardillocxx_amd import native, scenes

pytestmark = pytest.mark.gpu


def forward(ctx, prob, dt, p, v=None, minv=None, fext=None, mu=None):
    # f_gyro = -w x (I w) is recomputed on the device from the (perturbed) v and M^-1, so the
    # finite differences see the same dependence the adjoint differentiates
    ctx.set_bodies(prob.dof, prob.minv if minv is None else minv, prob.v if v is None else v,
                   prob.fext if fext is None else fext, None, prob.q)
    arr = list(prob.contact_arrays())
    if mu is not None:
        arr[2] = mu
    ctx.set_contacts(*arr)
    info = ctx.pj_solve(dt, 1.0, p)
    return ctx.velocities(), ctx.impulses(), info


def loss(v, imp, wv, wp):
    return float(wv @ v + (wp * imp).sum())


def params(**kw):
    p = native.default_params()
    for k, x in kw.items():
        setattr(p, k, x)
    return p


def directional_fd(f, x0, d, eps):
    return (f(x0 + eps * d) - f(x0 - eps * d)) / (2 * eps)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_implicit_adjoint_vs_finite_differences(ctx, seed):
    prob = scenes.random_problem(nb=12, nc=14, seed=200 + seed, warm=False)
    rng = np.random.default_rng(seed)
    dt = 1e-3
    p = params(tol_abs=1e-14, tol_rel=1e-14, max_iter=2_000_000, warmstart=0)
    wv = rng.normal(size=prob.nv)
    wp = rng.normal(size=(prob.nc, 3))
    v, imp, info = forward(ctx, prob, dt, p)
    g = ctx.pj_adjoint(dt, p, wv, wp, mode=native.CDX_ADJ_IMPLICIT, tol=1e-14)
    assert g["info"].mode_used == native.CDX_ADJ_IMPLICIT
    eps = 1e-7
    for name, base, grad in (("v", prob.v, g["v"]), ("minv", prob.minv, g["minv"]), ("fext", prob.fext, g["fext"]),
                             ("mu", prob.mu, g["mu"])):
        for _ in range(3):
            d = rng.normal(size=base.shape)
            if name == "mu":
                d = d * (prob.mu > 0)
            if name == "minv":  # keep the translational inverse mass isotropic
                d = d.reshape(-1) * base.reshape(-1)
                o = 0
                for nd in prob.dof:
                    d[o:o + 3] = d[o]
                    o += nd
            f = lambda x: loss(*forward(ctx, prob, dt, p, **{name: x})[:2], wv, wp)
            fd = directional_fd(f, base, d, eps * max(1.0, np.abs(base).max()) / max(1.0, np.abs(d).max()))
            an = float(grad @ d.reshape(-1))
            assert an == pytest.approx(fd, rel=2e-5, abs=1e-7 * max(1.0, abs(fd))), (name, an, fd)


def test_unrolled_matches_fixed_sweep_map_and_implicit(ctx):
    """Unrolled reverse mode is the exact derivative of the K-sweep map; at convergence it agrees
    with the implicit adjoint."""
    prob = scenes.random_problem(nb=10, nc=12, seed=77, warm=True)
    rng = np.random.default_rng(5)
    dt = 1e-3
    wv = rng.normal(size=prob.nv)
    wp = rng.normal(size=(prob.nc, 3))
    K = 40
    pK = params(max_iter=K, tol_abs=0.0, tol_rel=0.0)
    ctx.set_recording(True, K)
    forward(ctx, prob, dt, pK)
    g = ctx.pj_adjoint(dt, pK, wv, wp, mode=native.CDX_ADJ_UNROLLED)
    assert g["info"].mode_used == native.CDX_ADJ_UNROLLED and g["info"].iters == K
    for name, base, grad in (("v", prob.v, g["v"]), ("mu", prob.mu, g["mu"]), ("fext", prob.fext, g["fext"])):
        d = rng.normal(size=base.shape)
        if name == "mu":
            d = d * (prob.mu > 0)
        f = lambda x: loss(*forward(ctx, prob, dt, pK, **{name: x})[:2], wv, wp)
        fd = directional_fd(f, base, d, 1e-7)
        assert float(grad @ d.reshape(-1)) == pytest.approx(fd, rel=2e-5, abs=1e-8)
    # converged: unrolled (many sweeps) vs implicit
    Kc = 20000
    pc = params(max_iter=Kc, tol_abs=1e-14, tol_rel=1e-14)
    ctx.set_recording(True, Kc)
    forward(ctx, prob, dt, pc)
    gu = ctx.pj_adjoint(dt, pc, wv, wp, mode=native.CDX_ADJ_UNROLLED)
    gi = ctx.pj_adjoint(dt, pc, wv, wp, mode=native.CDX_ADJ_IMPLICIT, tol=1e-14)
    ctx.set_recording(False)
    for k in ("v", "fext", "mu"):
        np.testing.assert_allclose(gu[k], gi[k], rtol=1e-6, atol=1e-8 * np.abs(gi[k]).max())
