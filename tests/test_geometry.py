"""Geometry generators: sphere packing constraints and axis uniformity."""
import numpy as np
import pytest

from magpy_amd import geometry as G


@pytest.mark.parametrize("n,phi", [(2, 0.1), (33, 0.1), (64, 0.1), (512, 0.15)])
def test_random_cluster_constraints(n, phi):
    rng = np.random.default_rng(n)
    r = 10e-9
    p = G.random_cluster(n, r, phi, 2.2, rng)
    R = G.cluster_sphere_radius(n, r, phi)
    assert p.shape == (n, 3)
    assert np.all(np.linalg.norm(p, axis=1) <= R * (1 + 1e-12))
    d = np.linalg.norm(p[:, None, :] - p[None, :, :], axis=2)
    d[np.arange(n), np.arange(n)] = np.inf
    assert d.min() >= 2.2 * r * (1 - 1e-12)
    if n == 512:   # sphere radius r (512/0.15)^(1/3) = 150.6 nm
        assert abs(R - 150.6e-9) < 0.1e-9


def test_uniform_random_axes_are_uniform():
    from scipy import stats
    a = G.uniform_random_axes(20000, np.random.default_rng(1))
    assert np.allclose(np.linalg.norm(a, axis=1), 1.0)
    # each Cartesian component of a uniform point on S^2 is uniform on [-1, 1]
    for c in range(3):
        assert stats.kstest(a[:, c], "uniform", args=(-1, 2)).pvalue > 1e-3


def test_cluster_ensemble_layout():
    e = G.cluster_ensemble([2, 5, 3], 10e-9, 0.1, 2.2, np.random.default_rng(0))
    assert list(e["member_offsets"]) == [0, 2, 7, 10]
    assert e["location"].shape == (10, 3) and e["anisotropy_axis"].shape == (10, 3)
