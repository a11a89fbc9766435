"""Geometry of particle ensembles.

`packing` holds this build's seeded generators (uniform axes from a numpy
Generator, random sequential addition of particle centres in a sphere, whole
cluster ensembles); `coordinates` and `arkus` mirror owlas/magpy's
magpy/geometry package (chains, Arkus clusters, random rotations).
"""
from .packing import (cluster_ensemble, cluster_sphere_radius, random_cluster,  # noqa: F401
                      uniform_random_axes)
