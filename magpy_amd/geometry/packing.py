"""Seeded geometry and initial-condition generators.

`uniform_random_axes` draws points uniformly on the unit sphere with the
parameterisation of magpy/initial_conditions.py:4-40 (theta = 2 pi U1,
phi = arccos(1 - 2 U2)), from a numpy Generator instead of the global legacy
state.  `random_cluster` places particle centres in a sphere by random
sequential addition, the cluster geometry the interacting workloads use
(magpy ships no such generator).
"""
import numpy as np


def uniform_random_axes(n, rng):
    """n unit vectors uniform on the sphere, shape (n, 3)."""
    theta = 2.0 * np.pi * rng.uniform(size=n)
    phi = np.arccos(1 - 2.0 * rng.uniform(size=n))
    return np.stack([np.sin(phi) * np.cos(theta), np.sin(phi) * np.sin(theta), np.cos(phi)], 1)


def cluster_sphere_radius(n, radius, volume_fraction):
    """Radius of the sphere that holds n particles of `radius` at `volume_fraction`."""
    return radius * (n / volume_fraction) ** (1.0 / 3.0)


def random_cluster(n, radius, volume_fraction, min_distance_factor, rng, max_tries=100000):
    """Centres of n particles placed one at a time uniformly in a sphere of
    radius cluster_sphere_radius(n, radius, volume_fraction), each candidate
    rejected while it lies closer than min_distance_factor * radius to a
    placed centre.  Returns (n, 3) positions in metres, centred on the origin
    of the enclosing sphere."""
    R = cluster_sphere_radius(n, radius, volume_fraction)
    dmin2 = (min_distance_factor * radius) ** 2
    pts = np.zeros((n, 3))
    placed = 0
    tries = 0
    while placed < n:
        batch = 64
        u = rng.normal(size=(batch, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        cand = u * (R * rng.uniform(size=(batch, 1)) ** (1.0 / 3.0))
        for c in cand:
            tries += 1
            if placed == 0 or np.min(np.sum((pts[:placed] - c) ** 2, axis=1)) >= dmin2:
                pts[placed] = c
                placed += 1
                if placed == n:
                    break
        if tries > max_tries:
            raise RuntimeError("could not place %d particles (volume fraction too high?)" % n)
    return pts


def cluster_ensemble(sizes, radius, volume_fraction, min_distance_factor, rng, anisotropy=23e3):
    """Concatenated per-particle arrays of clusters with the given sizes:
    member_offsets, radius, anisotropy, easy axes (= initial magnetisation)
    and locations, as simulate_ensemble takes them."""
    sizes = np.asarray(sizes, dtype=np.int64)
    offs = np.zeros(sizes.size + 1, dtype=np.int64)
    offs[1:] = np.cumsum(sizes)
    P = int(offs[-1])
    loc = np.concatenate([random_cluster(int(n), radius, volume_fraction, min_distance_factor, rng)
                          for n in sizes])
    axes = uniform_random_axes(P, rng)
    return dict(member_offsets=offs, radius=np.full(P, radius), anisotropy=np.full(P, anisotropy),
                anisotropy_axis=axes, magnetisation_direction=axes.copy(), location=loc)
