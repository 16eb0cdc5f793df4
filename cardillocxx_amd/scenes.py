"""Scene builders and the contact problem container.

A :class:`ContactProblem` holds exactly what the reference's DynamicsAssembler hands a solver
(src/physics/assembly/dynamics_assembler.hpp:24-103): stacked body state and the contact buffer
in the layout of include/cdx/cdx.h.  Contact generation runs on the CPU (libcdx_scene.so).

Scenes (seeded, deterministic):
  * ``sphere_drop_grid``  -- 1000 spheres r=0.05 m, m=1 kg on a jittered 8x8 grid (pitch 0.12 m,
    16 layers) dropped into a 1x1x2 m box of static half-spaces (the plain CPU-runnable case).
  * ``sphere_pile``       -- 100k spheres, radius log-uniform in [0.02, 0.08] m, 1000 kg/m^3,
    random sequential placement in a 4x4x8 m box then a gravity settle; mu ~ U[0.3, 0.8] per body.
  * ``random_problem``    -- small random contact problems (test inputs).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np

from . import native

GRAVITY = np.array([0.0, 0.0, -9.81])


@dataclasses.dataclass
class ContactProblem:
    dof: np.ndarray  # int32 [nb]
    minv: np.ndarray  # [nv]
    v: np.ndarray  # [nv]
    fext: np.ndarray  # [nv]
    fgyro: np.ndarray  # [nv]
    q: np.ndarray  # [nq]
    body_a: np.ndarray  # int32 [nc]
    body_b: np.ndarray
    mu: np.ndarray  # [nc]
    frame: np.ndarray  # [nc, 9]
    angA: np.ndarray  # [nc, 9]
    angB: np.ndarray
    rest: np.ndarray  # [nc, 3]
    vbias: np.ndarray  # [nc, 3]
    warm: np.ndarray  # [nc, 3]
    point: np.ndarray | None = None  # [nc, 3] world contact points (warm-start matching)

    @property
    def nb(self):
        return len(self.dof)

    @property
    def nv(self):
        return int(self.dof.sum())

    @property
    def nc(self):
        return len(self.body_a)

    def contact_arrays(self):
        return (self.body_a, self.body_b, self.mu, self.frame, self.angA, self.angB, self.rest, self.vbias, self.warm)

    def body_arrays(self):
        return (self.dof, self.minv, self.v, self.fext, self.fgyro, self.q)

    def copy(self):
        return dataclasses.replace(self, **{f.name: (None if getattr(self, f.name) is None else getattr(self, f.name).copy())
                                            for f in dataclasses.fields(self)})


# --------------------------------------------------------------------------------------------
# bodies
# --------------------------------------------------------------------------------------------

def sphere_bodies(radius, mass, q, v=None):
    """Rigid spheres: M = [m,m,m, I,I,I] with I = 2/5 m r^2 (rigid_body_factory.cpp:67-69),
    f_ext = m g (world.cpp:113-135), v = [v_world; omega_body]."""
    radius = np.asarray(radius, dtype=np.float64)
    mass = np.broadcast_to(np.asarray(mass, dtype=np.float64), radius.shape)
    nb = len(radius)
    inertia = 0.4 * mass * radius * radius
    minv = np.empty((nb, 6))
    minv[:, :3] = (1.0 / mass)[:, None]
    minv[:, 3:] = (1.0 / inertia)[:, None]
    fext = np.zeros((nb, 6))
    fext[:, :3] = mass[:, None] * GRAVITY[None, :]
    vv = np.zeros((nb, 6)) if v is None else np.asarray(v, dtype=np.float64).reshape(nb, 6)
    return dict(dof=np.full(nb, 6, dtype=np.int32), minv=minv.reshape(-1), v=vv.reshape(-1).copy(),
                fext=fext.reshape(-1), fgyro=np.zeros(nb * 6), q=np.ascontiguousarray(q, dtype=np.float64).reshape(-1))


def gyro_force(dof, minv, v):
    """f_gyro = -omega x (I omega) (world.cpp:152-173) for every 6-dof body; zero elsewhere."""
    dof = np.asarray(dof)
    v = np.asarray(v, dtype=np.float64)
    minv = np.asarray(minv, dtype=np.float64)
    out = np.zeros_like(v)
    off = np.concatenate([[0], np.cumsum(dof)[:-1]]).astype(np.int64)
    rb = off[dof == 6]
    if len(rb):
        idx = rb[:, None] + np.arange(3, 6)[None, :]
        w = v[idx]
        mi = minv[idx]
        I = np.where(mi > 0, 1.0 / np.where(mi > 0, mi, 1.0), 0.0)
        out[idx] = -np.cross(w, I * w)
    return out


# --------------------------------------------------------------------------------------------
# contacts
# --------------------------------------------------------------------------------------------

def box_walls(lo, hi, mu_wall, top=False):
    """Static half-spaces bounding [lo, hi] (inward normals), open at the top unless `top`."""
    n, d = [], []
    n.append((0, 0, 1)); d.append(lo[2])
    n.append((1, 0, 0)); d.append(lo[0])
    n.append((-1, 0, 0)); d.append(-hi[0])
    n.append((0, 1, 0)); d.append(lo[1])
    n.append((0, -1, 0)); d.append(-hi[1])
    if top:
        n.append((0, 0, -1)); d.append(-hi[2])
    n = np.asarray(n, dtype=np.float64)
    return n, np.asarray(d, dtype=np.float64), np.full(len(d), float(mu_wall))


def sphere_contacts(q, radius, mu_body, walls, margin=0.0, friction=True):
    """CPU broad+narrow phase for spheres and static half-spaces (libcdx_scene)."""
    L = native.scene_lib()
    q = np.ascontiguousarray(q, dtype=np.float64)
    radius = np.ascontiguousarray(radius, dtype=np.float64)
    mu_body = np.ascontiguousarray(mu_body, dtype=np.float64)
    wn, wd, wmu = (np.ascontiguousarray(x, dtype=np.float64) for x in walls)
    nb = len(radius)
    cap = max(16, 8 * nb)
    while True:
        ia = np.empty(cap, np.int32); ib = np.empty(cap, np.int32)
        mu = np.empty(cap); frame = np.empty((cap, 9)); aA = np.empty((cap, 9)); aB = np.empty((cap, 9))
        pt = np.empty((cap, 3)); dep = np.empty(cap)
        n = L.cdx_sphere_contacts(nb, native.dptr(q), native.dptr(radius), native.dptr(mu_body), len(wd),
                                  native.dptr(wn.reshape(-1)), native.dptr(wd), native.dptr(wmu), margin, cap,
                                  native.iptr(ia), native.iptr(ib), native.dptr(mu), native.dptr(frame),
                                  native.dptr(aA), native.dptr(aB), native.dptr(pt), native.dptr(dep))
        if n < 0:
            raise RuntimeError("cdx_sphere_contacts failed")
        if n <= cap:
            break
        cap = n
    if not friction:
        mu = np.zeros(n)
    return dict(body_a=ia[:n].copy(), body_b=ib[:n].copy(), mu=mu[:n].copy(), frame=frame[:n].copy(),
                angA=aA[:n].copy(), angB=aB[:n].copy(), rest=np.zeros((n, 3)), vbias=np.zeros((n, 3)),
                warm=np.zeros((n, 3)), point=pt[:n].copy(), depth=dep[:n].copy())


def match_warmstart(prev, cur, max_dist=0.02):
    """ContactTracker greedy matching (contact_tracker.cpp:43-72); `prev` must carry `imp`."""
    L = native.scene_lib()
    warm = np.zeros((len(cur["body_a"]), 3))
    if len(prev["body_a"]) == 0 or len(cur["body_a"]) == 0:
        return warm
    args = [np.ascontiguousarray(prev["body_a"], np.int32), np.ascontiguousarray(prev["body_b"], np.int32),
            np.ascontiguousarray(prev["point"], np.float64), np.ascontiguousarray(prev["imp"], np.float64),
            np.ascontiguousarray(cur["body_a"], np.int32), np.ascontiguousarray(cur["body_b"], np.int32),
            np.ascontiguousarray(cur["point"], np.float64)]
    st = L.cdx_match_warmstart(len(args[0]), native.iptr(args[0]), native.iptr(args[1]), native.dptr(args[2]),
                               native.dptr(args[3]), len(args[4]), native.iptr(args[4]), native.iptr(args[5]),
                               native.dptr(args[6]), max_dist, native.dptr(warm))
    if st != 0:
        raise RuntimeError("cdx_match_warmstart failed")
    return warm


def make_problem(bodies: dict, contacts: dict) -> ContactProblem:
    keys = ["body_a", "body_b", "mu", "frame", "angA", "angB", "rest", "vbias", "warm"]
    return ContactProblem(**{k: bodies[k] for k in ["dof", "minv", "v", "fext", "fgyro", "q"]},
                          **{k: contacts[k] for k in keys}, point=contacts.get("point"))


# --------------------------------------------------------------------------------------------
# scenes
# --------------------------------------------------------------------------------------------

@dataclasses.dataclass
class SphereScene:
    q: np.ndarray  # [nb, 7]
    radius: np.ndarray
    mass: np.ndarray
    mu_body: np.ndarray
    walls: tuple
    margin: float = 0.0
    terrain: dict | None = None  # shared static heightfield (fractal_heightfield)
    disabled: tuple | None = None  # (a, b) body pairs that never collide (e.g. jointed links)

    def bodies(self, v=None):
        return sphere_bodies(self.radius, self.mass, self.q, v)

    def contacts(self, q=None, friction=True):
        qq = self.q if q is None else np.asarray(q).reshape(-1, 7)
        return sphere_contacts(qq, self.radius, self.mu_body, self.walls, self.margin, friction)

    def problem(self, friction=True):
        return make_problem(self.bodies(), self.contacts(friction=friction))


def sphere_drop_grid(n=1000, seed=0, mu=0.5):
    """1000 spheres (r=0.05, m=1) seeded on a jittered grid, dropped into a 1x1x2 m box."""
    L = native.scene_lib()
    q = np.empty((n, 7))
    L.cdx_scene_sphere_grid(n, 8, 8, 0.12, 0.06, 0.01, seed, native.dptr(q))
    radius = np.full(n, 0.05)
    return SphereScene(q=q, radius=radius, mass=np.ones(n), mu_body=np.full(n, mu),
                       walls=box_walls((-0.5, -0.5, 0.0), (0.5, 0.5, 2.0), mu))


def sphere_pile(n=100_000, seed=0, settle_steps=300, rmin=0.02, rmax=0.08, box=(4.0, 4.0, 8.0), density=1000.0,
                margin=None):
    """Settled polydisperse sphere pile (radii log-uniform, density 1000, mu ~ U[0.3, 0.8])."""
    L = native.scene_lib()
    q = np.empty((n, 7))
    r = np.empty(n)
    st = L.cdx_scene_sphere_pile(n, rmin, rmax, box[0], box[1], box[2], seed, settle_steps, native.dptr(q),
                                 native.dptr(r))
    if st != n:
        raise RuntimeError("cdx_scene_sphere_pile failed")
    rng = np.random.default_rng(seed + 1)
    mu_body = rng.uniform(0.3, 0.8, n)
    mass = density * 4.0 / 3.0 * math.pi * r ** 3
    return SphereScene(q=q, radius=r, mass=mass, mu_body=mu_body,
                       walls=box_walls((0.0, 0.0, 0.0), box, 0.8), margin=0.05 * rmin if margin is None else margin)


def _rand_rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]), q


def random_problem(nb=20, nc=60, seed=0, p_static=0.15, p_frictionless=0.2, p_point=0.1, rest=0.0, warm=True,
                   vscale=1.0):
    """Random well-posed contact problem: random masses/inertias, orthonormal contact frames,
    angular rows r_body x d_body of random lever arms, static sides, point masses, warm starts."""
    rng = np.random.default_rng(seed)
    dof = np.where(rng.random(nb) < p_point, 3, 6).astype(np.int32)
    minv, v, fext, q = [], [], [], []
    rots = []
    for d in dof:
        m = rng.uniform(0.5, 2.0)
        minv += [1 / m] * 3
        v += list(rng.normal(scale=vscale, size=3))
        fext += list(m * GRAVITY)
        R, qq = _rand_rot(rng)
        rots.append(R)
        if d == 6:
            I = rng.uniform(0.01, 0.1, 3)
            minv += list(1 / I)
            v += list(rng.normal(scale=vscale, size=3))
            fext += [0.0, 0.0, 0.0]
            q += list(rng.normal(size=3)) + list(qq)
        else:
            q += list(rng.normal(size=3))
    minv = np.array(minv); v = np.array(v); fext = np.array(fext); q = np.array(q)
    fgyro = gyro_force(dof, minv, v)
    ia = np.empty(nc, np.int32); ib = np.empty(nc, np.int32)
    frame = np.empty((nc, 9)); aA = np.zeros((nc, 9)); aB = np.zeros((nc, 9))
    for c in range(nc):
        # neighbours in index order keep the contact graph sparse (bounded degree), as in a pile
        a = int(rng.integers(0, nb))
        b = (a + 1 + int(rng.integers(0, min(3, nb - 1)))) % nb
        if a > b:
            a, b = b, a
        if rng.random() < p_static:
            if rng.random() < 0.5:
                a = -1
            else:
                b = -1
        ia[c], ib[c] = a, b
        n = rng.normal(size=3); n /= np.linalg.norm(n)
        t1 = np.cross(n, rng.normal(size=3)); t1 /= np.linalg.norm(t1)
        t2 = np.cross(n, t1)
        frame[c] = np.concatenate([n, t1, t2])
        for side, body, out in ((0, a, aA), (1, b, aB)):
            if body < 0 or dof[body] == 3:
                continue
            R = rots[body]
            r = rng.normal(scale=0.3, size=3)
            out[c] = np.concatenate([np.cross(r, R.T @ n), np.cross(r, R.T @ t1), np.cross(r, R.T @ t2)])
    mu = np.where(rng.random(nc) < p_frictionless, 0.0, rng.uniform(0.2, 0.9, nc))
    restitution = np.full((nc, 3), float(rest))
    vbias = np.where((ia < 0) | (ib < 0), 1, 0)[:, None] * rng.normal(scale=0.05, size=(nc, 3))
    w = np.zeros((nc, 3))
    if warm:
        w[:, 0] = np.abs(rng.normal(scale=0.1, size=nc))
        w[:, 1:] = rng.normal(scale=0.02, size=(nc, 2))
    return ContactProblem(dof=dof, minv=minv, v=v, fext=fext, fgyro=fgyro, q=q, body_a=ia, body_b=ib, mu=mu,
                          frame=frame, angA=aA, angB=aB, rest=restitution, vbias=vbias, warm=w)


def concat_problems(problems):
    """Concatenate independent contact problems into one batch: body ids of scene s are shifted
    by the bodies before it; returns (problem, body_off, contact_off)."""
    boff, coff = [0], [0]
    parts = {k: [] for k in ("dof", "minv", "v", "fext", "fgyro", "q", "body_a", "body_b", "mu", "frame", "angA",
                             "angB", "rest", "vbias", "warm")}
    for p in problems:
        shift = boff[-1]
        for k in ("dof", "minv", "v", "fext", "fgyro", "q", "mu", "frame", "angA", "angB", "rest", "vbias", "warm"):
            parts[k].append(getattr(p, k))
        parts["body_a"].append(np.where(p.body_a >= 0, p.body_a + shift, -1).astype(np.int32))
        parts["body_b"].append(np.where(p.body_b >= 0, p.body_b + shift, -1).astype(np.int32))
        boff.append(boff[-1] + p.nb)
        coff.append(coff[-1] + p.nc)
    cat = {k: np.concatenate(v) if v else np.zeros(0) for k, v in parts.items()}
    cat["dof"] = cat["dof"].astype(np.int32)
    for k in ("frame", "angA", "angB"):
        cat[k] = cat[k].reshape(-1, 9)
    for k in ("rest", "vbias", "warm"):
        cat[k] = cat[k].reshape(-1, 3)
    return ContactProblem(**cat), np.array(boff, np.int32), np.array(coff, np.int32)


def split_velocities(problems, v):
    out, o = [], 0
    for p in problems:
        out.append(v[o:o + p.nv])
        o += p.nv
    return out


# --------------------------------------------------------------------------------------------
# spheres and boxes in many independent scenes
# --------------------------------------------------------------------------------------------

SPHERE, BOX, CAPSULE = 0, 1, 2


def capsule_unit_inertia(radius, half_length):
    """Inertia per unit mass of a capsule about its centre (axis = body z): the reference divides
    Coal's Capsule::computeMomentofInertia() (unit density) by the volume
    (rigid_body_factory.cpp:76-79; Coal v3.0.4, the cylinder plus two hemispheres by the
    parallel-axis theorem)."""
    r = np.asarray(radius, dtype=np.float64)
    h = np.asarray(half_length, dtype=np.float64)
    v_cyl = math.pi * r * r * (2.0 * h)
    v_sph = 4.0 / 3.0 * math.pi * r ** 3
    ix = v_cyl * (h * h / 3.0 + r * r / 4.0) + v_sph * (0.4 * r * r + h * h + 0.75 * r * h)
    iz = (0.5 * v_cyl + 0.4 * v_sph) * r * r
    vol = v_cyl + v_sph
    return np.stack([ix / vol, ix / vol, iz / vol], axis=-1)


def capsule_volume(radius, half_length):
    return math.pi * radius * radius * (2.0 * half_length) + 4.0 / 3.0 * math.pi * radius ** 3


def shape_bodies(shape, dims, mass, q, v=None):
    """Rigid spheres/boxes/capsules: sphere I = 2/5 m r^2, box I = m (h_j^2 + h_k^2) / 3, capsule
    from capsule_unit_inertia (computeUnitInertia, rigid_body_factory.cpp:59-91)."""
    shape = np.asarray(shape)
    dims = np.asarray(dims, dtype=np.float64).reshape(-1, 3)
    mass = np.asarray(mass, dtype=np.float64)
    nb = len(shape)
    I = np.empty((nb, 3))
    sph = shape == SPHERE
    box = shape == BOX
    cap = shape == CAPSULE
    I[sph] = (0.4 * mass[sph] * dims[sph, 0] ** 2)[:, None]
    h2 = dims ** 2
    I[box, 0] = mass[box] * (h2[box, 1] + h2[box, 2]) / 3.0
    I[box, 1] = mass[box] * (h2[box, 0] + h2[box, 2]) / 3.0
    I[box, 2] = mass[box] * (h2[box, 0] + h2[box, 1]) / 3.0
    if cap.any():
        I[cap] = mass[cap, None] * capsule_unit_inertia(dims[cap, 0], dims[cap, 1])
    minv = np.concatenate([np.repeat((1.0 / mass)[:, None], 3, axis=1), 1.0 / I], axis=1)
    fext = np.zeros((nb, 6))
    fext[:, :3] = mass[:, None] * GRAVITY[None, :]
    vv = np.zeros((nb, 6)) if v is None else np.asarray(v, dtype=np.float64).reshape(nb, 6)
    dof = np.full(nb, 6, dtype=np.int32)
    fgyro = gyro_force(dof, minv.reshape(-1), vv.reshape(-1))
    return dict(dof=dof, minv=minv.reshape(-1), v=vv.reshape(-1).copy(), fext=fext.reshape(-1), fgyro=fgyro,
                q=np.ascontiguousarray(q, dtype=np.float64).reshape(-1))


def _terrain_struct(terrain):
    if terrain is None:
        return None, None
    h = np.ascontiguousarray(terrain["h"], dtype=np.float64).reshape(-1)
    t = native.Heightfield(int(terrain["nx"]), int(terrain["ny"]), float(terrain["x0"]), float(terrain["y0"]),
                           float(terrain["dx"]), float(terrain["dy"]), native.dptr(h), float(terrain["mu"]))
    return t, h


def shape_contacts(body_off, q, shape, dims, mu_body, walls, margin=0.0, friction=True, terrain=None, disabled=None):
    """CPU contacts for spheres/boxes/capsules in independent scenes, optionally on a shared
    heightfield, skipping the body pairs in `disabled` ((a, b) arrays) (libcdx_scene
    cdx_shape_contacts_hf)."""
    L = native.scene_lib()
    hf, hf_keep = _terrain_struct(terrain)
    da, db = (np.zeros(0, np.int32), np.zeros(0, np.int32)) if disabled is None else \
        (np.ascontiguousarray(disabled[0], np.int32), np.ascontiguousarray(disabled[1], np.int32))
    body_off = np.ascontiguousarray(body_off, np.int32)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
    shape = np.ascontiguousarray(shape, np.int32)
    dims = np.ascontiguousarray(dims, dtype=np.float64).reshape(-1)
    mu_body = np.ascontiguousarray(mu_body, dtype=np.float64)
    wn, wd, wmu = (np.ascontiguousarray(x, dtype=np.float64) for x in walls)
    ns = len(body_off) - 1
    cap = max(64, 8 * len(shape))
    coff = np.zeros(ns + 1, np.int32)
    while True:
        ia = np.empty(cap, np.int32); ib = np.empty(cap, np.int32)
        mu = np.empty(cap); frame = np.empty((cap, 9)); aA = np.empty((cap, 9)); aB = np.empty((cap, 9))
        pt = np.empty((cap, 3)); dep = np.empty(cap)
        n = L.cdx_shape_contacts_hf(ns, native.iptr(body_off), native.dptr(q), native.iptr(shape), native.dptr(dims),
                                    native.dptr(mu_body), len(wd), native.dptr(wn.reshape(-1)), native.dptr(wd),
                                    native.dptr(wmu), None if hf is None else C.byref(hf), len(da), native.iptr(da),
                                    native.iptr(db), margin, cap, native.iptr(ia),
                                    native.iptr(ib), native.dptr(mu), native.dptr(frame), native.dptr(aA), native.dptr(aB),
                                    native.dptr(pt), native.dptr(dep), native.iptr(coff))
        if n < 0:
            raise RuntimeError("cdx_shape_contacts_hf failed")
        if n <= cap:
            break
        cap = n
    if not friction:
        mu = np.zeros(n)
    return dict(body_a=ia[:n].copy(), body_b=ib[:n].copy(), mu=mu[:n].copy(), frame=frame[:n].copy(),
                angA=aA[:n].copy(), angB=aB[:n].copy(), rest=np.zeros((n, 3)), vbias=np.zeros((n, 3)),
                warm=np.zeros((n, 3)), point=pt[:n].copy(), depth=dep[:n].copy(), contact_off=coff)


@dataclasses.dataclass
class ShapeBatch:
    """Independent scenes of spheres and boxes; scene s owns bodies [body_off[s], body_off[s+1])."""
    body_off: np.ndarray
    shape: np.ndarray
    dims: np.ndarray  # [nb, 3] sphere radius in column 0, box half extents
    mass: np.ndarray
    mu_body: np.ndarray
    q: np.ndarray  # [nb, 7]
    v: np.ndarray  # [nb, 6]
    walls: tuple
    margin: float = 0.0
    terrain: dict | None = None  # shared static heightfield (fractal_heightfield)
    disabled: tuple | None = None  # (a, b) body pairs that never collide (e.g. jointed links)

    @property
    def nscenes(self):
        return len(self.body_off) - 1

    def bodies(self):
        return shape_bodies(self.shape, self.dims, self.mass, self.q, self.v)

    def scene(self, s):
        """Scene s on its own (a one-scene batch with the same walls and margin)."""
        b0, b1 = int(self.body_off[s]), int(self.body_off[s + 1])
        return ShapeBatch(body_off=np.array([0, b1 - b0], np.int32), shape=self.shape[b0:b1].copy(),
                          dims=self.dims[b0:b1].copy(), mass=self.mass[b0:b1].copy(), mu_body=self.mu_body[b0:b1].copy(),
                          q=self.q[b0:b1].copy(), v=self.v[b0:b1].copy(), walls=self.walls, margin=self.margin,
                          terrain=self.terrain, disabled=None if self.disabled is None else _scene_pairs(self.disabled, b0, b1))

    def contacts(self, q=None, friction=True):
        return shape_contacts(self.body_off, self.q if q is None else q, self.shape, self.dims, self.mu_body, self.walls,
                              self.margin, friction, self.terrain, self.disabled)

    def problem(self, friction=True):
        c = self.contacts(friction=friction)
        return make_problem(self.bodies(), c), c["contact_off"]


def _scene_pairs(pairs, b0, b1):
    a, b = np.asarray(pairs[0]), np.asarray(pairs[1])
    keep = (a >= b0) & (a < b1) & (b >= b0) & (b < b1)
    return (a[keep] - b0).astype(np.int32), (b[keep] - b0).astype(np.int32)


def rsa(body_off, bound, lo, hi, seed):
    """Random sequential placement of bounding spheres per scene (libcdx_scene cdx_scene_rsa)."""
    body_off = np.ascontiguousarray(body_off, np.int32)
    bound = np.ascontiguousarray(bound, np.float64)
    pos = np.zeros((len(bound), 3))
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray(hi, np.float64)
    native.scene_lib().cdx_scene_rsa(len(body_off) - 1, native.iptr(body_off), native.dptr(bound), native.dptr(lo),
                                     native.dptr(hi), seed, native.dptr(pos))
    return pos


def batched_drop(nscenes=4096, seed=0, nmin=8, nmax=512, density=1000.0, mu=0.5):
    """Independent scenes, bodies per scene log-uniform in [nmin, nmax], spheres and boxes 50/50
    (box edges U[0.05, 0.2] m, sphere radius of the equal-volume sphere), placed at random without
    overlap in a unit box and released from rest."""
    rng = np.random.default_rng(seed)
    counts = np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax + 1), nscenes))).astype(int)
    counts = np.clip(counts, nmin, nmax)
    body_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    nb = int(body_off[-1])
    shape = np.where(rng.random(nb) < 0.5, SPHERE, BOX).astype(np.int32)
    edges = rng.uniform(0.05, 0.2, (nb, 3))
    vol = edges.prod(axis=1)
    dims = 0.5 * edges
    rad = (3.0 * vol / (4.0 * math.pi)) ** (1.0 / 3.0)
    dims[shape == SPHERE] = rad[shape == SPHERE, None] * np.array([1.0, 0.0, 0.0])
    mass = density * vol
    bound = np.where(shape == SPHERE, rad, np.linalg.norm(dims, axis=1))
    q = np.zeros((nb, 7))
    q[:, :3] = rsa(body_off, bound, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), seed)
    qq = rng.normal(size=(nb, 4))
    q[:, 3:] = qq / np.linalg.norm(qq, axis=1, keepdims=True)
    return ShapeBatch(body_off=body_off, shape=shape, dims=dims, mass=mass, mu_body=np.full(nb, mu), q=q,
                      v=np.zeros((nb, 6)), walls=box_walls((0, 0, 0), (1.0, 1.0, 1.0), mu), margin=1e-3)


def boxes_on_plane(nscenes=256, per_scene=64, seed=0, edge=0.1, mass=1.0, vstd=0.5, spacing=0.25):
    """Boxes (edge 0.1 m, 1 kg) resting on a plane on an 8x8 grid, horizontal initial velocities
    N(0, vstd^2), friction mu ~ U[0.2, 0.9] per box."""
    rng = np.random.default_rng(seed)
    nb = nscenes * per_scene
    side = int(math.ceil(math.sqrt(per_scene)))
    gx, gy = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    grid = np.stack([gx.ravel(), gy.ravel()], 1)[:per_scene] * spacing
    q = np.zeros((nb, 7))
    q[:, 0:2] = np.tile(grid, (nscenes, 1))
    q[:, 2] = 0.5 * edge
    q[:, 6] = 1.0
    v = np.zeros((nb, 6))
    v[:, :2] = rng.normal(scale=vstd, size=(nb, 2))
    walls = (np.array([[0.0, 0.0, 1.0]]), np.array([0.0]), np.array([1.0]))
    return ShapeBatch(body_off=(np.arange(nscenes + 1) * per_scene).astype(np.int32), shape=np.full(nb, BOX, np.int32),
                      dims=np.full((nb, 3), 0.5 * edge), mass=np.full(nb, float(mass)), mu_body=rng.uniform(0.2, 0.9, nb),
                      q=q, v=v, walls=walls, margin=1e-4)


def fractal_heightfield(n=256, extent=6.4, amp=0.1, octaves=6, seed=0, mu=0.5):
    """Seeded fractal value noise on an n x n grid over [0, extent]^2: octave k is a random
    lattice of 2^(k+2) + 1 values bilinearly interpolated, weighted 0.5^k; scaled so that
    max |h| = amp."""
    rng = np.random.default_rng(seed)
    u = np.linspace(0.0, 1.0, n)
    h = np.zeros((n, n))
    for k in range(octaves):
        m = 2 ** (k + 2) + 1
        lat = rng.uniform(-1.0, 1.0, (m, m))
        x = u * (m - 1)
        i = np.minimum(np.floor(x).astype(int), m - 2)
        f = x - i
        rows = lat[i] * (1 - f)[:, None] + lat[i + 1] * f[:, None]            # interpolate along the first axis
        h += 0.5 ** k * (rows[:, i] * (1 - f)[None, :] + rows[:, i + 1] * f[None, :])
    h *= amp / max(np.abs(h).max(), 1e-300)
    dx = extent / (n - 1)
    # h[i + n j] at (x0 + i dx, y0 + j dy): row index j (y), column index i (x)
    return dict(nx=n, ny=n, x0=0.0, y0=0.0, dx=dx, dy=dx, h=np.ascontiguousarray(h.T.reshape(-1)), mu=mu)


def terrain_height(terrain, x, y):
    """Height of the triangulated heightfield at (x, y) (same split as the contact generator)."""
    fx = (np.asarray(x) - terrain["x0"]) / terrain["dx"]
    fy = (np.asarray(y) - terrain["y0"]) / terrain["dy"]
    nx, ny = terrain["nx"], terrain["ny"]
    i = np.clip(np.floor(fx).astype(int), 0, nx - 2)
    j = np.clip(np.floor(fy).astype(int), 0, ny - 2)
    a, b = fx - i, fy - j
    H = terrain["h"]
    h00, h10, h01, h11 = H[i + nx * j], H[i + 1 + nx * j], H[i + nx * (j + 1)], H[i + 1 + nx * (j + 1)]
    lower = a + b <= 1.0
    return np.where(lower, h00 + a * (h10 - h00) + b * (h01 - h00),
                    h11 + (1 - a) * (h01 - h11) + (1 - b) * (h10 - h11))


def capsule_chains(nchains=2048, links=32, radius=0.02, length=0.1, density=1000.0, terrain=None, drop=0.15, seed=0,
                   mu=0.5, local=False):
    """Chains of `links` capsules (total length `length` including the caps, radius `radius`),
    one chain per scene, laid out tip to tip along a random horizontal direction `drop` above the
    highest terrain point (with `local`, above the highest point under that chain), over a
    fractal heightfield shared by all scenes.  Returns the batch and the spherical joints between
    consecutive links: (body_a, body_b, anchor_a, anchor_b) with the anchors at the touching tips
    in body coordinates."""
    terrain = terrain if terrain is not None else fractal_heightfield(seed=seed)
    rng = np.random.default_rng(seed + 1)
    hl = 0.5 * length - radius
    nb = nchains * links
    extent = terrain["dx"] * (terrain["nx"] - 1)
    span = links * length
    q = np.zeros((nb, 7))
    top = float(np.max(terrain["h"])) + radius + drop
    for s in range(nchains):
        ang = rng.uniform(0, 2 * math.pi)
        d = np.array([math.cos(ang), math.sin(ang), 0.0])
        c = rng.uniform(0.5 * span + 0.05, extent - 0.5 * span - 0.05, 2) if extent > span + 0.1 else \
            np.full(2, 0.5 * extent)
        z = top
        if local:
            sx = c[0] + np.linspace(-0.5 * span, 0.5 * span, 4 * links + 1) * d[0]
            sy = c[1] + np.linspace(-0.5 * span, 0.5 * span, 4 * links + 1) * d[1]
            z = float(np.max(terrain_height(terrain, sx, sy))) + radius + drop
        start = np.array([c[0], c[1], z]) - 0.5 * span * d
        # body z axis along d: rotate e_z by 90 degrees about e_z x d
        axis = np.cross([0.0, 0.0, 1.0], d)
        axis /= np.linalg.norm(axis)
        quat = np.concatenate([axis * math.sin(math.pi / 4), [math.cos(math.pi / 4)]])
        for k in range(links):
            b = s * links + k
            q[b, :3] = start + (k + 0.5) * length * d
            q[b, 3:] = quat
    dims = np.zeros((nb, 3))
    dims[:, 0] = radius
    dims[:, 1] = hl
    mass = np.full(nb, density * capsule_volume(radius, hl))
    batch = ShapeBatch(body_off=(np.arange(nchains + 1) * links).astype(np.int32), shape=np.full(nb, CAPSULE, np.int32),
                       dims=dims, mass=mass, mu_body=np.full(nb, mu), q=q, v=np.zeros((nb, 6)),
                       walls=(np.zeros((0, 3)), np.zeros(0), np.zeros(0)), margin=1e-3, terrain=terrain)
    ja = np.concatenate([np.arange(s * links, (s + 1) * links - 1) for s in range(nchains)]).astype(np.int32)
    jb = ja + 1
    batch.disabled = (ja, jb)  # jointed links do not collide (beam_factory.cpp:235)
    anchor_a = np.tile([0.0, 0.0, hl + radius], (len(ja), 1))
    anchor_b = np.tile([0.0, 0.0, -(hl + radius)], (len(ja), 1))
    return batch, dict(body_a=ja, body_b=jb, anchor_a=anchor_a, anchor_b=anchor_b)
