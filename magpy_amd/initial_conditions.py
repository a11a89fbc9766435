"""Initial conditions (magpy/initial_conditions.py:1-56): uniformly random
axes and rotations drawn from numpy's global legacy state in the reference's
order, so np.random.seed(s) reproduces the reference's draws.

Quaternions are (w, x, y, z) arrays, the convention of the transforms3d
functions the reference calls (axangle2quat, rotate_vector); both are
written here from their definitions.
"""
import numpy as np


def random_point_on_unit_sphere():
    """theta = 2 pi U1, phi = arccos(1 - 2 U2) (initial_conditions.py:4-16)."""
    theta = 2.0 * np.pi * np.random.rand()
    phi = np.arccos(1 - 2.0 * np.random.rand())
    return np.array([np.sin(phi) * np.cos(theta), np.sin(phi) * np.sin(theta), np.cos(phi)])


def uniform_random_axes(N):
    """N unit vectors uniform on the sphere, shape (N, 3) (initial_conditions.py:19-35)."""
    return np.array([random_point_on_unit_sphere() for _ in range(N)])


def axangle2quat(vector, theta):
    """Unit quaternion of a rotation by theta about `vector` (normalised here):
    (cos(theta/2), sin(theta/2) * v / |v|)."""
    v = np.asarray(vector, dtype=np.float64)
    v = v / np.sqrt(np.dot(v, v))
    s = np.sin(theta / 2.0)
    return np.array([np.cos(theta / 2.0), s * v[0], s * v[1], s * v[2]])


def _qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def rotate_vector(v, q):
    """v rotated by the quaternion q: the vector part of q (0, v) q*."""
    qc = np.array([q[0], -q[1], -q[2], -q[3]])
    return _qmul(_qmul(q, np.concatenate([[0.0], np.asarray(v, dtype=np.float64)])), qc)[1:]


def random_quaternion():
    """Rotation about a uniformly random axis by 2 pi times a standard normal
    angle (initial_conditions.py:38-56)."""
    return axangle2quat(random_point_on_unit_sphere(), 2.0 * np.pi * np.random.randn())
