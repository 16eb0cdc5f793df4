"""Batched solve: every scene of a one-launch batch equals a standalone solve of that scene."""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import native, scenes

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    den = max(np.abs(b).max(initial=0.0), 1e-300)
    return np.abs(np.asarray(a) - np.asarray(b)).max(initial=0.0) / den


def make_batch(n, seed):
    rng = np.random.default_rng(seed)
    probs = []
    for s in range(n):
        nb = int(np.exp(rng.uniform(np.log(4), np.log(120))))
        nc = int(rng.integers(0, int(1.5 * nb)))
        probs.append(scenes.random_problem(nb=nb, nc=nc, seed=1000 * seed + s, warm=bool(s % 2)))
    return probs


def upload_batch(ctx, probs):
    cat, boff, coff = scenes.concat_problems(probs)
    ctx.set_bodies(cat.dof, cat.minv, cat.v, cat.fext, cat.fgyro, cat.q)
    ctx.set_contacts(*cat.contact_arrays())
    ctx.set_scenes(boff, coff)
    return cat, boff, coff


@pytest.mark.parametrize("launch", [(256, 1), (64, 1), (512, 4), (128, 2)])
def test_batch_matches_standalone_oracle(ctx, launch):
    probs = make_batch(40, seed=1)
    cat, boff, coff = upload_batch(ctx, probs)
    p = native.default_params()
    p.tol_abs = p.tol_rel = 1e-8
    ctx.batch_set_launch(*launch)
    iters, res = ctx.batch_pj_solve(1e-3, 1.0, p)
    v = ctx.velocities()
    imp = ctx.impulses()
    for s, prob in enumerate(probs):
        st, v_ref, imp_ref, r = oracle.pj_solve(prob, 1e-3, oracle.params(tol_abs=1e-8, tol_rel=1e-8))
        assert iters[s] == r.iters, s
        vs = v[cat.dof[:boff[s]].sum():cat.dof[:boff[s + 1]].sum()]
        assert rel(vs, v_ref) <= TOL
        if prob.nc:
            assert rel(imp[coff[s]:coff[s + 1]], imp_ref) <= TOL


def test_batch_lds_scenes_equal_global_scenes(ctx):
    """One scene per workgroup keeps the scene's velocities and side-impulse records in LDS when
    every scene fits; two scenes per workgroup keep them in global memory.  The per-contact and
    per-body arithmetic is the same, so are iterations, velocities and impulses, bit for bit; the
    residual is summed over the group's waves and may differ in the last bits."""
    probs = make_batch(24, seed=5)
    p = native.default_params()
    p.tol_abs = p.tol_rel = 1e-7
    out = []
    try:
        for launch in ((128, 1), (128, 2), (512, 1)):
            upload_batch(ctx, probs)  # the same warm start for every launch
            ctx.batch_set_launch(*launch)
            iters, res = ctx.batch_pj_solve(1e-3, 1.0, p)
            out.append((iters, res, ctx.velocities(), ctx.impulses()))
    finally:
        ctx.batch_set_launch(256, 1)
    for o in out[1:]:
        assert np.array_equal(out[0][0], o[0])
        assert np.allclose(out[0][1], o[1], rtol=1e-12, atol=0, equal_nan=True)
        assert np.array_equal(out[0][2], o[2]) and np.array_equal(out[0][3], o[3])


def test_batch_scene_too_large_for_lds(ctx):
    """A scene whose velocities and records do not fit in the LDS budget sends the whole batch
    through the global-memory kernel, also at one scene per workgroup."""
    probs = [scenes.random_problem(nb=1400, nc=2000, seed=7), scenes.random_problem(nb=10, nc=12, seed=8)]
    cat, boff, coff = upload_batch(ctx, probs)
    p = native.default_params()
    p.tol_abs = p.tol_rel = 1e-8
    ctx.batch_set_launch(256, 1)
    iters, res = ctx.batch_pj_solve(1e-3, 1.0, p)
    v = ctx.velocities()
    for s, prob in enumerate(probs):
        st, v_ref, imp_ref, r = oracle.pj_solve(prob, 1e-3, oracle.params(tol_abs=1e-8, tol_rel=1e-8))
        assert iters[s] == r.iters, s
        assert rel(v[cat.dof[:boff[s]].sum():cat.dof[:boff[s + 1]].sum()], v_ref) <= TOL


def test_batch_fixed_sweeps_equals_grid_solve(ctx):
    """Same per-contact arithmetic as the grid-wide solve (which sums the side impulses in a
    different order): a one-scene batch agrees with it to rounding."""
    prob = scenes.random_problem(nb=300, nc=450, seed=5)
    p = native.default_params()
    p.max_iter, p.tol_abs, p.tol_rel = 77, 0.0, 0.0
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    ctx.set_contacts(*prob.contact_arrays())
    ctx.pj_solve(1e-3, 1.0, p)
    v1, i1 = ctx.velocities(), ctx.impulses()
    upload_batch(ctx, [prob])
    iters, _ = ctx.batch_pj_solve(1e-3, 1.0, p)
    assert iters[0] == 77
    assert np.abs(ctx.velocities() - v1).max() <= 1e-12 * np.abs(v1).max()
    assert np.abs(ctx.impulses() - i1).max() <= 1e-12 * np.abs(i1).max()


def test_batch_autotune_and_sphere_scenes(ctx):
    probs = []
    for s in range(16):
        pile = scenes.sphere_pile(n=60 + 20 * s, settle_steps=60, box=(0.5, 0.5, 1.0), seed=s)
        probs.append(pile.problem())
    upload_batch(ctx, probs)
    p = native.default_params()
    p.tol_abs = p.tol_rel = 1e-6
    block, groups, ms = ctx.batch_autotune(1e-3, 1.0, p)
    assert block in (64, 128, 256, 512) and ms > 0
    iters, res = ctx.batch_pj_solve(1e-3, 1.0, p)
    v = ctx.velocities()
    o = 0
    for s, prob in enumerate(probs):
        st, v_ref, imp_ref, r = oracle.pj_solve(prob, 1e-3, oracle.params(tol_abs=1e-6, tol_rel=1e-6))
        assert iters[s] == r.iters
        assert rel(v[o:o + prob.nv], v_ref) <= TOL
        o += prob.nv


def test_scene_validation(ctx):
    probs = make_batch(3, seed=2)
    cat, boff, coff = scenes.concat_problems(probs)
    ctx.set_bodies(cat.dof, cat.minv, cat.v, cat.fext, cat.fgyro, cat.q)
    ctx.set_contacts(*cat.contact_arrays())
    bad = boff.copy()
    bad[-1] -= 1  # does not cover every body
    with pytest.raises(native.CdxError):
        ctx.set_scenes(bad, coff)
    # a contact of scene 0 moved into scene 1's range
    c0 = int(np.argmax(cat.body_a >= 0))
    s0 = int(np.searchsorted(coff, c0, side="right") - 1)
    moved = coff.copy()
    moved[s0 + 1] = c0  # contact c0 now belongs to scene s0+1 but references a body of scene s0
    if 0 < c0 < coff[-1]:
        with pytest.raises(native.CdxError):
            ctx.set_scenes(boff, np.sort(moved))
    with pytest.raises(native.CdxError):
        ctx.batch_set_launch(256, 3)


CHEB_CASES = [
    dict(nb=20, nc=60, seed=2),
    dict(nb=50, nc=200, seed=3, p_point=0.3),
    dict(nb=200, nc=300, seed=4, rest=0.4),
    dict(nb=3, nc=8, seed=6, p_static=0.3, p_point=1.0, warm=False),  # full column rank: rho < 0.995
]


@pytest.mark.parametrize("launch", [(256, 1), (512, 4)])
@pytest.mark.parametrize("sweeps", [1, 2, 9, 64])
def test_batch_chebyshev_fixed_sweeps_matches_oracle(ctx, launch, sweeps):
    """Per-scene Chebyshev loop (chebyshev_loop with estimateSpectralRadius,
    projected_jacobi.cpp:145-228): every scene's own spectral radius and w recurrence, equal to
    the oracle's standalone Chebyshev solve of that scene after the same sweeps."""
    probs = [scenes.random_problem(**c) for c in CHEB_CASES] + make_batch(12, seed=3)
    cat, boff, coff = upload_batch(ctx, probs)
    p = native.default_params()
    p.accel = native.CDX_ACCEL_CHEBYSHEV
    p.max_iter, p.tol_abs, p.tol_rel = sweeps, 0.0, 0.0
    try:
        ctx.batch_set_launch(*launch)
        iters, res = ctx.batch_pj_solve(1e-3, 1.0, p)
    finally:
        ctx.batch_set_launch(256, 1)
    v, imp = ctx.velocities(), ctx.impulses()
    for s, prob in enumerate(probs):
        st, v_ref, imp_ref, r = oracle.pj_solve(prob, 1e-3, oracle.params(chebyshev=1, max_iter=sweeps, tol_abs=0.0,
                                                                          tol_rel=0.0))
        assert st == 0 and iters[s] == r.iters, s
        vs = v[cat.dof[:boff[s]].sum():cat.dof[:boff[s + 1]].sum()]
        assert rel(vs, v_ref) <= TOL, s
        if prob.nc:
            assert rel(imp[coff[s]:coff[s + 1]], imp_ref) <= TOL, s


@pytest.mark.parametrize("tol", [1e-5, 1e-9])
def test_batch_chebyshev_to_tolerance(ctx, tol):
    """Scenes on which the Chebyshev loop converges: the same iteration count as the oracle and
    the converged sweep's input returned, per scene."""
    probs = [scenes.random_problem(**CHEB_CASES[0]), scenes.random_problem(**CHEB_CASES[3])]
    probs += [scenes.sphere_pile(n=40 + 20 * k, settle_steps=60, box=(0.5, 0.5, 1.0), seed=k).problem() for k in range(4)]
    cat, boff, coff = upload_batch(ctx, probs)
    p = native.default_params()
    p.accel = native.CDX_ACCEL_CHEBYSHEV
    p.tol_abs = p.tol_rel = tol
    iters, res = ctx.batch_pj_solve(1e-3, 1.0, p)
    v, imp = ctx.velocities(), ctx.impulses()
    for s, prob in enumerate(probs):
        st, v_ref, imp_ref, r = oracle.pj_solve(prob, 1e-3, oracle.params(chebyshev=1, tol_abs=tol, tol_rel=tol))
        assert st == 0 and iters[s] == r.iters and r.residual <= 1, s
        vs = v[cat.dof[:boff[s]].sum():cat.dof[:boff[s + 1]].sum()]
        assert rel(vs, v_ref) <= TOL, s
        if prob.nc:
            assert rel(imp[coff[s]:coff[s + 1]], imp_ref) <= TOL, s


def test_batch_refuses_other_accelerations(ctx):
    probs = make_batch(3, seed=2)
    upload_batch(ctx, probs)
    p = native.default_params()
    for mode in (native.CDX_ACCEL_NESTEROV, native.CDX_ACCEL_ANDERSON):
        p.accel = mode
        with pytest.raises(native.CdxError):
            ctx.batch_pj_solve(1e-3, 1.0, p)


def block_batch():
    """Random scenes, boxes resting on a plane (four-corner manifolds: coupled frictional blocks)
    and a small sphere pile."""
    probs = [scenes.random_problem(**c) for c in CHEB_CASES[:3]] + make_batch(8, seed=11)
    plane = scenes.boxes_on_plane(nscenes=2, per_scene=16, seed=4)
    probs += [plane.scene(s).problem()[0] for s in range(plane.nscenes)]
    probs.append(scenes.sphere_pile(n=60, settle_steps=60, box=(0.5, 0.5, 1.0), seed=2).problem())
    return probs


@pytest.mark.parametrize("launch", [(256, 1), (512, 4)])
@pytest.mark.parametrize("accel", ["none", "chebyshev"])
@pytest.mark.parametrize("sweeps", [1, 7, 40])
def test_batch_block_preconditioner_fixed_sweeps_matches_oracle(ctx, launch, accel, sweeps):
    """Per-contact inverse Delassus blocks in the batched solve (pj.rdiag_true_delassus,
    projected_jacobi.cpp:71-98, 330-348), standard and Chebyshev loops, LDS-resident (256, 1) and
    global (512, 4) kernels: every scene equals the oracle's standalone block solve after the same
    sweeps."""
    probs = block_batch()
    cat, boff, coff = upload_batch(ctx, probs)
    p = native.default_params()
    p.block_precond = 1
    p.accel = native.CDX_ACCEL_CHEBYSHEV if accel == "chebyshev" else native.CDX_ACCEL_NONE
    p.max_iter, p.tol_abs, p.tol_rel = sweeps, 0.0, 0.0
    try:
        ctx.batch_set_launch(*launch)
        iters, res = ctx.batch_pj_solve(1e-3, 1.0, p)
    finally:
        ctx.batch_set_launch(256, 1)
    v, imp = ctx.velocities(), ctx.impulses()
    for s, prob in enumerate(probs):
        st, v_ref, imp_ref, r = oracle.pj_solve(prob, 1e-3, oracle.params(block_precond=1, chebyshev=int(accel == "chebyshev"),
                                                                          max_iter=sweeps, tol_abs=0.0, tol_rel=0.0))
        assert st == 0 and iters[s] == r.iters, s
        vs = v[cat.dof[:boff[s]].sum():cat.dof[:boff[s + 1]].sum()]
        assert rel(vs, v_ref) <= TOL, s
        if prob.nc:
            assert rel(imp[coff[s]:coff[s + 1]], imp_ref) <= TOL, s


@pytest.mark.parametrize("launch", [(256, 1), (128, 2)])
def test_batch_block_preconditioner_to_tolerance(ctx, launch):
    """Block mode to tol 1e-8 per scene: the oracle's iteration count, velocities and impulses."""
    probs = block_batch()
    cat, boff, coff = upload_batch(ctx, probs)
    p = native.default_params()
    p.block_precond = 1
    p.tol_abs = p.tol_rel = 1e-8
    try:
        ctx.batch_set_launch(*launch)
        iters, res = ctx.batch_pj_solve(1e-3, 1.0, p)
    finally:
        ctx.batch_set_launch(256, 1)
    v, imp = ctx.velocities(), ctx.impulses()
    for s, prob in enumerate(probs):
        st, v_ref, imp_ref, r = oracle.pj_solve(prob, 1e-3, oracle.params(block_precond=1, tol_abs=1e-8, tol_rel=1e-8))
        assert st == 0 and iters[s] == r.iters, s
        vs = v[cat.dof[:boff[s]].sum():cat.dof[:boff[s + 1]].sum()]
        assert rel(vs, v_ref) <= TOL, s
        if prob.nc:
            assert rel(imp[coff[s]:coff[s + 1]], imp_ref) <= TOL, s
