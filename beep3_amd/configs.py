"""Seeded synthetic workloads (meshes, charges, poses) for the five measured
configurations of the BEM solve.  Rules: near pairs use 6-point quadrature on
the column patch (72 points per patch) with centroid collocation on the row
patch; far pairs use the 1-point centroid interaction.
"""
from __future__ import annotations

import numpy as np

from .beep import BEEP
from .scene import Mesh

QUAD, QUAL = 6, 0  # "1 quadrature pt far / 6 pts near"


def born_ion(subdivisions=4, nbsize=1500, **kw):
    """Icosphere r = 2 A (2,562 vertices at level 4), +1 e at the centre,
    eps_in 2, eps_out 80, kappa 0.125."""
    m = Mesh.icosphere(subdivisions, 2.0).set_charges([[0.0, 0.0, 0.0, 1.0]])
    b = BEEP(80.0, 0.125, QUAD, QUAL, nbsize, **kw)
    b.insert_mesh_instance(b.load_library_mesh(m), None, None, 2.0)
    return b


def off_centre(subdivisions=6, n_charges=200, seed=2, nbsize=1500, **kw):
    """Icosphere r = 20 A (40,962 vertices at level 6), 200 charges uniform in
    the ball r < 17 A with |q| ~ U(0.25, 1.0) and random sign; eps_in 4,
    eps_out 80, kappa 0.1."""
    m = Mesh.icosphere(subdivisions, 20.0).charges_ball(n_charges, 17.0, 0.25, 1.0, seed)
    b = BEEP(80.0, 0.1, QUAD, QUAL, nbsize, **kw)
    b.insert_mesh_instance(b.load_library_mesh(m), None, None, 4.0)
    return b


def globular_mesh(subdivisions=7, shape_seed=3, n_charges=2000, charge_seed=3):
    """Star-shaped surface r = 25 A (1 + sum_{l=2..8} a_lm Y_lm), a_lm ~
    N(0, 0.04/l) (standard deviation), on an icosphere (163,842 vertices at
    level 7); charges from a 3.8 A random walk kept 2 A inside, q ~ N(0, 0.4)."""
    m = Mesh.star(subdivisions, 25.0, 8, 0.04, shape_seed)
    return m.charges_walk(n_charges, 3.8, 2.0, 0.4, charge_seed)


def globular(subdivisions=7, shape_seed=3, n_charges=2000, charge_seed=3, nbsize=1500, **kw):
    """eps_in 2, eps_out 80, kappa 0.125."""
    m = globular_mesh(subdivisions, shape_seed, n_charges, charge_seed)
    b = BEEP(80.0, 0.125, QUAD, QUAL, nbsize, **kw)
    b.insert_mesh_instance(b.load_library_mesh(m), None, None, 2.0)
    return b


def random_pose(rng: np.random.Generator, sep_lo=45.0, sep_hi=80.0):
    """Random unit quaternion (w, x, y, z) and a centre offset of uniform
    length in [sep_lo, sep_hi] along a random direction."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    return q, d * rng.uniform(sep_lo, sep_hi)


def pose_pair(mesh: Mesh, q, offset, nbsize=1500, **kw):
    """Two instances of one mesh: the first at the origin, the second rotated
    by q with its centre at offset (eps_in 2, eps_out 80, kappa 0.125)."""
    b = BEEP(80.0, 0.125, QUAD, QUAL, nbsize, **kw)
    lib_id = b.load_library_mesh(mesh)
    b.insert_mesh_instance(lib_id, [0.0, 0.0, 0.0], None, 2.0)
    b.insert_mesh_instance(lib_id, offset, q, 2.0)
    return b


# ---- serving stream: Poisson arrivals of single-molecule solves ----
SERVING_LEVELS = (4, 5, 6, 7)            # 2,562 / 10,242 / 40,962 / 163,842 patches
SERVING_WEIGHTS = (0.4, 0.3, 0.2, 0.1)


def serving_schedule(rate=50.0, duration=6.0, seed=11, shapes=2):
    """Arrival times of a Poisson stream (rate per second over duration
    seconds) with, per request, the icosphere level (weights 0.4 / 0.3 / 0.2
    / 0.1 over 2,562 .. 163,842 patches), which of `shapes` star shapes of
    that level, the charge count (uniform 50..2,000) and the charge seed."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        lvl = int(rng.choice(SERVING_LEVELS, p=SERVING_WEIGHTS))
        out.append({"t": t, "level": lvl, "shape": int(rng.integers(shapes)),
                    "n_charges": int(rng.integers(50, 2001)), "charge_seed": int(rng.integers(1 << 30))})
    return out


def serving_meshes(shapes=2, levels=SERVING_LEVELS):
    """Star surfaces (r = 25 A, the 163,842-patch globular generator at each
    level) with shape seeds 100 + k; charges are set per request."""
    return {(lvl, k): Mesh.star(lvl, 25.0, 8, 0.04, 100 + k) for lvl in levels for k in range(shapes)}


def serving_request(meshes, req, nbsize=1500):
    """Scene of one request: its own copy of the surface with a 3.8 A
    random-walk chain of n_charges charges, q ~ N(0, 0.4) (eps_in 2, eps_out
    80, kappa 0.125)."""
    from .scene import Scene
    m = meshes[(req["level"], req["shape"])].clone()
    m.charges_walk(req["n_charges"], 3.8, 2.0, 0.4, req["charge_seed"])
    return Scene(80.0, 0.125, QUAD, QUAL, nbsize).add_instance(m, 2.0), m


# ---- batched poses: two instances of one surface in many rigid poses ----
def pose_mesh(subdivisions=6, shape_seed=7, n_charges=2000, charge_seed=7):
    """The 163,842-patch globular generator at 40,962 patches with a_lm seed 7
    (r = 25 A star surface, 3.8 A random-walk charges, q ~ N(0, 0.4))."""
    return globular_mesh(subdivisions, shape_seed, n_charges, charge_seed)


def pose_scenes(mesh: Mesh, n_poses=256, seed=7, nbsize=1500):
    """Scenes of n_poses seeded rigid poses: instance 1 at the origin,
    instance 2 rotated by a random unit quaternion with its centre at a
    uniform 45-80 A distance in a random direction (eps_in 2, eps_out 80,
    kappa 0.125).  All scenes share the one mesh."""
    from .scene import Scene
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_poses):
        q, off = random_pose(rng)
        sc = Scene(80.0, 0.125, QUAD, QUAL, nbsize)
        sc.add_instance(mesh, 2.0, quat=None, offset=[0.0, 0.0, 0.0])
        sc.add_instance(mesh, 2.0, quat=q, offset=off)
        out.append(sc)
    return out
