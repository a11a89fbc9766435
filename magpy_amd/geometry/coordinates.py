"""Particle coordinates of chains and Arkus clusters, and random rotations
(magpy/geometry/coordinates.py:1-123).  Random draws use numpy's global
legacy state in the reference's order, so np.random.seed(s) reproduces the
reference's coordinates."""
import numpy as np

from ..initial_conditions import axangle2quat, rotate_vector
from .arkus import ARKUS


def arkus_cluster_coordinates(n_particles, configuration_id, R):
    """ARKUS[n_particles][configuration_id] scaled by the distance R
    (coordinates.py:7-28)."""
    return ARKUS[n_particles][configuration_id] * R


def arkus_cluster_random_configuration_id(n_particles):
    """A uniformly drawn configuration id of an n_particles Arkus cluster
    (coordinates.py:31-47)."""
    return np.random.randint(len(ARKUS[n_particles]))


def arkus_random_cluster_coordinates(n_particles, R):
    """Coordinates of a randomly chosen configuration (coordinates.py:50-70)."""
    return arkus_cluster_coordinates(n_particles,
                                     arkus_cluster_random_configuration_id(n_particles), R)


def chain_coordinates(n_particles, R, direction=np.array([0, 0, 1])):
    """n_particles centres R apart on a straight line from the origin along
    the unit vector of `direction` (coordinates.py:72-104)."""
    d = np.asarray(direction, dtype=np.float64)
    unit = d / np.linalg.norm(d)
    return R * np.outer(np.arange(n_particles), unit)


def random_rotate_coordinates(coords):
    """All coordinates rotated about one random axis by a random angle
    (coordinates.py:106-123): axis uniform on the sphere (two uniform
    draws), angle 2 pi times a standard normal draw."""
    theta = 2.0 * np.pi * np.random.rand()
    phi = np.arccos(1 - 2.0 * np.random.rand())
    axis = np.array([np.sin(phi) * np.cos(theta), np.sin(phi) * np.sin(theta), np.cos(phi)])
    q = axangle2quat(axis, 2.0 * np.pi * np.random.randn())
    return np.array([rotate_vector(c, q) for c in np.atleast_2d(coords)])
