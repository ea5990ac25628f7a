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
