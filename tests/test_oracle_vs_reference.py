"""The oracle's restatement against owlas/magpy's own building blocks compiled
from source (oracle/_ref/libmagpy_ref.so: lib/llg.cpp, lib/field.cpp,
lib/rng.cpp, lib/distances.cpp, lib/trap.cpp), on seeded random inputs.

Both sides are built without FMA contraction, so agreement is bit-exact.
The library is built from /root/reference when that tree is present; where it
is absent and no prebuilt library travelled with the tree, these tests skip.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

L = O.lib()
ptr, f64 = O.ptr, O.f64


def _ref():
    R = O.ref()
    if R is None and os.path.isdir(os.path.join(O.REFERENCE_TREE, "lib")):
        O.build(ref=True)
        R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built and the reference tree is absent")
    return R


@pytest.fixture(scope="module")
def R():
    return _ref()


RNG = np.random.default_rng(20240607)


def _vec(n):
    return f64(RNG.normal(size=n))


def test_drift_diffusion_and_jacobians(R):
    for _ in range(500):
        m, h, hj = _vec(3), _vec(3), _vec(9)
        a, sr = RNG.uniform(0, 2), RNG.uniform(0, 2)
        for fo, fr, args, n in [
            (L.orc_llg_drift, R.ref_llg_drift, (ptr(m), a, ptr(h)), 3),
            (L.orc_llg_diffusion, R.ref_llg_diffusion, (ptr(m), sr, a), 9),
            (L.orc_llg_drift_jacobian, R.ref_llg_drift_jacobian, (ptr(m), a, ptr(h), ptr(hj)), 9),
            (L.orc_llg_diffusion_jacobian, R.ref_llg_diffusion_jacobian, (ptr(m), sr, a), 27),
        ]:
            o, r = np.zeros(n), np.zeros(n)
            fo(ptr(o), *args)
            fr(ptr(r), *args)
            assert np.array_equal(o, r)


@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_multi_particle_layouts(R, n):
    n3 = 3 * n
    for _ in range(20):
        m, h, axes = _vec(n3), _vec(n3), _vec(n3)
        srs, al, kr = f64(RNG.uniform(0, 1, n)), f64(RNG.uniform(0, 1, n)), f64(RNG.uniform(0, 2, n))
        # diffusion (dense, zeros written)
        o, r = np.full(n3 * n3, 5.0), np.full(n3 * n3, 5.0)
        L.orc_llg_multi_diffusion(ptr(o), ptr(m), ptr(srs), ptr(al), n)
        R.ref_llg_multi_diffusion(ptr(r), ptr(m), ptr(srs), ptr(al), n)
        assert np.array_equal(o, r)
        # field Jacobian, then the drift quasi-Jacobian reading it
        hjo, hjr = np.zeros(n3 * n3), np.zeros(n3 * n3)
        L.orc_field_multi_add_uniaxial_jacobian(ptr(hjo), ptr(axes), ptr(kr), n)
        R.ref_field_multi_add_uniaxial_jacobian(ptr(hjr), ptr(axes), ptr(kr), n)
        assert np.array_equal(hjo, hjr)
        o, r = np.zeros(n3 * n3), np.zeros(n3 * n3)
        L.orc_llg_multi_drift_quasijacobian(ptr(o), ptr(m), ptr(al), ptr(h), ptr(hjo), n)
        R.ref_llg_multi_drift_quasijacobian(ptr(r), ptr(m), ptr(al), ptr(h), ptr(hjr), n)
        assert np.array_equal(o, r)
        o, r = np.zeros(n3 ** 3), np.zeros(n3 ** 3)
        L.orc_llg_multi_diffusion_jacobian(ptr(o), ptr(m), ptr(srs), ptr(al), n)
        R.ref_llg_multi_diffusion_jacobian(ptr(r), ptr(m), ptr(srs), ptr(al), n)
        assert np.array_equal(o, r)
        # anisotropy field added on top of an existing field
        base = _vec(n3)
        o, r = base.copy(), base.copy()
        L.orc_field_multi_add_uniaxial(ptr(o), ptr(m), ptr(axes), ptr(kr), n)
        R.ref_field_multi_add_uniaxial(ptr(r), ptr(m), ptr(axes), ptr(kr), n)
        assert np.array_equal(o, r)


@pytest.mark.parametrize("n", [2, 3, 7, 16])
def test_dipolar_and_distances(R, n):
    pts = f64(RNG.uniform(-5e-8, 5e-8, size=3 * n))
    uo, mo = np.zeros(n * n * 3), np.zeros(n * n)
    vecs = np.zeros(n * n * 3)
    L.orc_pair_distance_vectors(ptr(vecs), ptr(pts), n)
    L.orc_pair_distance_magnitudes(ptr(mo), ptr(vecs), n)
    L.orc_pair_unit_vectors(ptr(uo), ptr(vecs), ptr(mo), n)
    ur, mr = np.zeros(n * n * 3), np.zeros(n * n)
    R.ref_distances(ptr(pts), n, ptr(ur), ptr(mr))
    assert np.array_equal(mo, mr)
    off = ~np.eye(n, dtype=bool)
    assert np.array_equal(uo.reshape(n, n, 3)[off], ur.reshape(n, n, 3)[off])
    assert np.all(np.isnan(uo.reshape(n, n, 3)[~off]))  # 0/0 on the diagonal, never read
    cubes = f64(np.where(off, (mo.reshape(n, n) / 1e-8) ** 3, 0.0).reshape(-1))
    m, v = _vec(3 * n), f64(RNG.uniform(0.5, 1.5, n))
    for ms, kav in [(4e5, 2.3e4), (1 / np.sqrt(1.25663706e-6), 0.5)]:
        o, r = np.zeros(3 * n), np.zeros(3 * n)
        units = f64(np.nan_to_num(uo))
        L.orc_field_multi_add_dipolar(ptr(o), ms, kav, ptr(v), ptr(m), ptr(units), ptr(cubes), n)
        R.ref_field_multi_add_dipolar(ptr(r), ms, kav, ptr(v), ptr(m), ptr(units), ptr(cubes), n)
        assert np.array_equal(o, r)
        assert L.orc_dipolar_prefactor(ms, kav) == R.ref_dipolar_prefactor(ms, kav)


def test_field_functions(R):
    for _ in range(2000):
        t, h, f = RNG.uniform(0, 1e3), RNG.normal(), RNG.uniform(0, 1e-2)
        assert L.orc_field_sinusoidal(t, h, f) == R.ref_field_sinusoidal(t, h, f)
        assert L.orc_field_square(t, h, f) == R.ref_field_square(t, h, f)


def test_rng_mt_norm_stream(R):
    import ctypes as C
    g = C.create_string_buffer(4096)
    L.orc_mtnorm_init.argtypes = [C.c_void_p, C.c_uint64, C.c_double]
    L.orc_mtnorm_get.restype = C.c_double
    L.orc_mtnorm_get.argtypes = [C.c_void_p]
    for seed, sd in [(1, 1.0), (1001, 1.0), (2 ** 31 - 2, 0.3), (123456789, 2.0)]:
        r = np.zeros(20000)
        R.ref_rng_mtnorm(seed, sd, ptr(r), r.size)
        L.orc_mtnorm_init(g, seed, sd)
        o = np.array([L.orc_mtnorm_get(g) for _ in range(r.size)])
        assert np.array_equal(o, r)


def test_trapezoidal(R):
    for n in (2, 3, 10, 1000):
        x = f64(np.sort(RNG.uniform(0, 1, n)))
        y = _vec(n)
        assert L.orc_trapezoidal(ptr(x), ptr(y), n) == R.ref_trapezoidal(ptr(x), ptr(y), n)
