"""magpy/geometry and magpy/initial_conditions equivalents.

The known answers are the reference's own (test/test_geometries.py:8-53):
the chain, the exact coordinates of Arkus cluster (7, 3) and the validity
and mean of random configuration ids.  Rotations are checked against what
the mathematics fixes: a quarter turn about z maps x to y, rotations keep
all pairwise distances, and q and -q rotate alike."""
import numpy as np

from magpy_amd import initial_conditions as IC
from magpy_amd.geometry.arkus import ARKUS
from magpy_amd.geometry.coordinates import (arkus_cluster_coordinates,
                                            arkus_cluster_random_configuration_id,
                                            arkus_random_cluster_coordinates, chain_coordinates,
                                            random_rotate_coordinates)


def test_chains():
    coords = chain_coordinates(3, 2.5, [1, 0, 0])
    assert np.all(coords == np.array([[0, 0, 0], [2.5, 0, 0], [5.0, 0, 0]]))
    assert np.allclose(chain_coordinates(2, 3.0, direction=[3, 4, 0]), [[0, 0, 0], [1.8, 2.4, 0]])


def test_arkus_select():
    coords = arkus_cluster_coordinates(7, 3, 5e-9)
    ans = np.array([
        [0.0, 0.0, 0.0],
        [0.0, 1.618033988749913, 0.0],
        [0.951056516295152, 0.309016994374951, 0.0],
        [-0.587785252292505, 0.809016994374957, -0.000000000000003],
        [0.951056516295155, 1.309016994374955, -0.000000000000003],
        [0.262865556059561, 0.809016994374957, -0.525731112119131],
        [0.262865556059561, 0.809016994374957, 0.525731112119127]
    ]) * 5e-9
    assert np.all(coords == ans)


def test_arkus_table_shape():
    assert sorted(ARKUS) == list(range(1, 10))
    assert [len(ARKUS[n]) for n in range(1, 10)] == [1, 1, 1, 1, 1, 2, 5, 13, 0]
    for n in range(1, 9):
        for c in ARKUS[n].values():
            assert c.shape == (n, 3)


def test_random_config_ids_are_valid():
    np.random.seed(0)
    for n_particles in range(1, 8):
        ids = [arkus_cluster_random_configuration_id(n_particles) for _ in range(2000)]
        assert np.all([i in ARKUS[n_particles] for i in ids])


def test_random_config_id_distribution_6_particle():
    np.random.seed(1)
    ids = [arkus_cluster_random_configuration_id(6) for _ in range(100000)]
    assert np.abs(np.mean(ids) - 0.5) < 1e-2


def test_quaternion_rotation_known_answers():
    q = IC.axangle2quat([0, 0, 5.0], np.pi / 2)
    assert np.allclose(IC.rotate_vector([1.0, 0, 0], q), [0, 1.0, 0], atol=1e-15)
    assert np.allclose(IC.rotate_vector([1.0, 2.0, 3.0], -q), IC.rotate_vector([1.0, 2.0, 3.0], q))
    q2 = IC.axangle2quat([1.0, 1.0, 1.0], 2 * np.pi / 3)      # cyclic permutation x->y->z
    assert np.allclose(IC.rotate_vector([1.0, 0, 0], q2), [0, 1.0, 0], atol=1e-15)


def test_random_rotation_keeps_distances():
    np.random.seed(2)
    c = arkus_random_cluster_coordinates(8, 1.0)
    r = random_rotate_coordinates(c)
    d0 = np.linalg.norm(c[:, None] - c[None], axis=2)
    d1 = np.linalg.norm(r[:, None] - r[None], axis=2)
    assert np.allclose(d0, d1, atol=1e-12)
    assert not np.allclose(c, r)


def test_uniform_random_axes():
    np.random.seed(3)
    ax = IC.uniform_random_axes(20000)
    assert np.allclose(np.linalg.norm(ax, axis=1), 1.0)
    # uniform on the sphere: each coordinate is uniform on [-1, 1]
    for d in range(3):
        h, _ = np.histogram(ax[:, d], bins=10, range=(-1, 1))
        assert np.all(np.abs(h - 2000) < 5 * np.sqrt(2000))
    q = IC.random_quaternion()
    assert abs(np.linalg.norm(q) - 1) < 1e-12
