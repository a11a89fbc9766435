"""The oracle against owlas/magpy's own known-answer tests (test/tests.cpp),
published algorithm vectors (Philox4x32-10, mt19937_64) and closed forms.

Every expected value here comes from the reference's test file (cited by
line), from a published known-answer vector, or from a closed-form solution;
none is recomputed with the oracle's own formula.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O

L = O.lib()
ptr, f64 = O.ptr, O.f64


def dbl_eq(a, b, ulps=4):
    """gtest EXPECT_DOUBLE_EQ: within 4 ULPs."""
    if a == b:
        return True
    return abs(a - b) <= ulps * np.spacing(max(abs(a), abs(b)))


# --------------------------------------------------------------- lib/llg.cpp
def test_llg_drift_known_answer():  # test/tests.cpp:23-34
    d = np.zeros(3)
    L.orc_llg_drift(ptr(d), ptr(f64([2, 3, 4])), 3.0, ptr(f64([1, 2, 3])))
    assert list(d) == [-34, -4, 20]


def test_llg_diffusion_known_answer():  # test/tests.cpp:36-52
    d = np.zeros(9)
    L.orc_llg_diffusion(ptr(d), ptr(f64([2, 3, 4])), 2.0, 3.0)
    assert list(d) == [150, -28, -54, -44, 120, -68, -42, -76, 78]


def test_llg_multi_diffusion_layout():  # test/tests.cpp:54-88
    d = np.full(36, 7.0)
    L.orc_llg_multi_diffusion(ptr(d), ptr(f64([2, 3, 4, 2, 3, 4])), ptr(f64([2, 2])),
                              ptr(f64([3, 3])), 2)
    d = d.reshape(6, 6)
    blk = np.array([[150, -28, -54], [-44, 120, -68], [-42, -76, 78]], dtype=float)
    for i in range(6):
        for j in range(6):
            if i // 3 != j // 3:
                assert d[i, j] == 0
    assert np.array_equal(d[:3, :3], blk) and np.array_equal(d[3:, 3:], blk)


def test_llg_drift_jacobian_known_answer():  # test/tests.cpp:94-111 (Mathematica values)
    j = np.zeros(9)
    L.orc_llg_drift_jacobian(ptr(j), ptr(f64([1.0, 1.2, 2.2])), 0.1,
                             ptr(f64([6.7, 13.4, 10.05])),
                             ptr(f64([1.0, 2.0, 1.5, 2.0, 4.0, 3.0, 1.5, 3.0, 2.25])))
    exp = [-1.161, -4.466, 19.33, 11.878, -2.977, None, -14.046, 3.8, -4.051]
    for k, e in enumerate(exp):
        if e is None:
            assert abs(j[k] - (-2.082)) <= 1e-10
        else:
            assert dbl_eq(j[k], e), (k, j[k], e)


def test_drift_jacobian_matches_finite_difference_of_drift():
    # The Jacobian with a field linear in m (h = K m, hj = K) against central
    # differences of the drift itself: catches a transposed or dropped term.
    rng = np.random.default_rng(3)
    for _ in range(20):
        m = rng.normal(size=3)
        K = rng.normal(size=(3, 3))
        a = rng.uniform(0.01, 1.0)

        def drift(mm):
            d = np.zeros(3)
            L.orc_llg_drift(ptr(d), ptr(f64(mm)), a, ptr(f64(K @ mm)))
            return d

        j = np.zeros(9)
        L.orc_llg_drift_jacobian(ptr(j), ptr(f64(m)), a, ptr(f64(K @ m)), ptr(f64(K)))
        # layout: jac[3*r + c] = d drift_r / d m_c (row-major Jacobian)
        fd = np.zeros((3, 3))
        eps = 1e-6
        for i in range(3):
            e = np.zeros(3)
            e[i] = eps
            fd[i] = (drift(m + e) - drift(m - e)) / (2 * eps)
        assert np.allclose(j.reshape(3, 3), fd.T, rtol=1e-6, atol=1e-7)


def test_diffusion_jacobian_matches_finite_difference():
    rng = np.random.default_rng(4)
    for _ in range(10):
        m = rng.normal(size=3)
        sr, a = rng.uniform(0.1, 2), rng.uniform(0.01, 1)

        def diff(mm):
            d = np.zeros(9)
            L.orc_llg_diffusion(ptr(d), ptr(f64(mm)), sr, a)
            return d.reshape(3, 3)

        jc = np.zeros(27)
        L.orc_llg_diffusion_jacobian(ptr(jc), ptr(f64(m)), sr, a)
        jc = jc.reshape(3, 3, 3)  # [x][y][z] = d B[x][y] / d m[z]
        eps = 1e-6
        # The reference writes entry 4 (dB01/dm1) as -a sr m2 and entry 25
        # (dB22/dm1) as 2 a sr m2 (lib/llg.cpp:128,156); the exact
        # derivatives are -a sr m0 and 2 a sr m1.  Those two entries only
        # steer Newton's path, and are pinned bit-exact against the
        # reference build in test_oracle_vs_reference.py instead.
        quirk = {(0, 1, 1), (2, 2, 1)}
        for z in range(3):
            e = np.zeros(3)
            e[z] = eps
            fd = (diff(m + e) - diff(m - e)) / (2 * eps)
            for x in range(3):
                for y in range(3):
                    if (x, y, z) in quirk:
                        assert not np.isclose(jc[x, y, z], fd[x, y], rtol=1e-6, atol=1e-8)
                    else:
                        assert np.isclose(jc[x, y, z], fd[x, y], rtol=1e-6, atol=1e-8)


# --------------------------------------------------------------- lib/field.cpp
def test_uniaxial_known_answer():  # test/tests.cpp:117-126
    h = np.zeros(3)
    L.orc_field_uniaxial(ptr(h), ptr(f64([1.0, 1.2, 2.2])), ptr(f64([1.0, 2.0, 1.5])))
    for a, e in zip(h, [6.7, 13.4, 10.05]):
        assert dbl_eq(a, e)


def test_uniaxial_jacobian_known_answer():  # test/tests.cpp:132-146
    j = np.zeros(9)
    L.orc_field_uniaxial_jacobian(ptr(j), ptr(f64([1.0, 2.0, 1.5])))
    assert list(j) == [1.0, 2.0, 1.5, 2.0, 4.0, 3.0, 1.5, 3.0, 2.25]


def test_multi_uniaxial_jacobian_known_answer():  # test/tests.cpp:842-904
    jac = np.zeros(81)
    L.orc_field_multi_add_uniaxial_jacobian(ptr(jac), ptr(f64([1, 2, 3, 4, 5, 6, 4, 2, 3])),
                                            ptr(f64([3.2, 5.2, 1.2])), 3)
    r = jac.reshape(9, 9)
    for i in range(9):
        for j in range(9):
            if i // 3 != j // 3:
                assert r[i, j] == 0.0
    exp = {(0, 0): 3.2, (0, 1): 6.4, (0, 2): 9.6, (1, 0): 6.4, (1, 1): 12.8, (1, 2): 19.2,
           (2, 0): 9.6, (2, 1): 19.2, (2, 2): 28.8,
           (3, 3): 83.2, (3, 4): 104, (3, 5): 124.8, (4, 3): 104, (4, 4): 130, (4, 5): 156,
           (5, 3): 124.8, (5, 4): 156, (5, 5): 187.2,
           (6, 6): 19.2, (6, 7): 9.6, (6, 8): 14.4, (7, 6): 9.6, (7, 7): 4.8, (7, 8): 7.2,
           (8, 6): 14.4, (8, 7): 7.2, (8, 8): 10.8}
    for (i, j), e in exp.items():
        assert dbl_eq(r[i, j], e), (i, j, r[i, j], e)


def test_applied_field_multi_add():  # test/tests.cpp:826-840 (h(t)=2t at t=3)
    h = np.full(6, 2.0)
    L.orc_field_multi_add_applied_z(ptr(h), 2.0 * 3.0, 2)
    assert list(h) == [2.0, 2.0, 8.0, 2.0, 2.0, 8.0]


def test_dipolar_prefactor_known_answer():  # test/tests.cpp:804-809
    ms = 1 / math.sqrt(O.C.c_double.in_dll(L, "ORC_MU0").value)
    assert dbl_eq(L.orc_dipolar_prefactor(ms, 0.5), 0.07957747154594767280)


def test_dipolar_p2p_known_answer():  # test/tests.cpp:811-824
    out = f64([1.0, 1.5, 2.0])
    L.orc_dipolar_add_p2p_term(ptr(out), 2.3, 1.2, ptr(f64([0.7, 0.2, 0.4])),
                               ptr(f64([1.2, 1.1, 1.3])), 3.7)
    for a, e in zip(out, [36.37323333333333333333, 37.05761666666666666667,
                          42.86218333333333333333]):
        assert dbl_eq(a, e)


def test_dipolar_two_particles_known_answer():  # test/tests.cpp:622-643
    mu0 = O.C.c_double.in_dll(L, "ORC_MU0").value
    field = np.zeros(6)
    s3 = math.sqrt(3.0) / 2.0
    dists = f64([0, 0, 0, -0.5, 0, s3, 0.5, 0, -s3, 0, 0, 0])
    L.orc_field_multi_add_dipolar(ptr(field), 1.0 / math.sqrt(mu0), 0.5,
                                  ptr(f64([15. / 16, 17. / 16])), ptr(f64([1, 0, 0, 0, 0, 1])),
                                  ptr(dists), ptr(f64([0, 1, 1, 0])), 2)
    exp = [-0.10983505338481014402, 0, 0.10568882939696175316, -0.01865096989358148646, 0,
           -0.09691328239836190239]
    for a, e in zip(field, exp):
        assert dbl_eq(a, e), (a, e)


@pytest.mark.parametrize("r2,k1,k2", [(2.5e-9, 6.33e4, 6.33e4), (3.0e-9, 5e4, 7e4)])
def test_dipolar_aligned_pair_closed_form(r2, k1, k2):
    # test/tests.cpp:645-802: point-dipole field of the partner, H = V Ms/(4 pi R^3)(3(m.r)r - m)
    mu0 = O.C.c_double.in_dll(L, "ORC_MU0").value
    r1 = 2.5e-9
    v1, v2 = 4. / 3 * math.pi * r1 ** 3, 4. / 3 * math.pi * r2 ** 3
    vav, kav, ms, R = 0.5 * (v1 + v2), 0.5 * (k1 + k2), 400e3, 2.917e-9
    rr = R / vav ** (1. / 3)
    dists = f64([0, 0, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0])
    cubes = f64([0, rr ** 3, rr ** 3, 0])
    hk = 2 * kav / mu0 / ms
    n1 = math.sqrt(0.14)
    n2 = math.sqrt(0.9)
    for mag in ([1, 0, 0, 1, 0, 0], [0, 0, 1, 0, 0, 1],
                [0.1 / n1, 0.2 / n1, 0.3 / n1, -0.8 / n2, 0.5 / n2, -0.1 / n2]):
        field = np.zeros(6)
        L.orc_field_multi_add_dipolar(ptr(field), ms, kav, ptr(f64([v1 / vav, v2 / vav])),
                                      ptr(f64(mag)), ptr(dists), ptr(cubes), 2)
        c1 = v2 * ms / (4.0 * math.pi * R ** 3)
        c2 = v1 * ms / (4.0 * math.pi * R ** 3)
        exp = [c1 * -mag[3] / hk, c1 * -mag[4] / hk, c1 * 2 * mag[5] / hk,
               c2 * -mag[0] / hk, c2 * -mag[1] / hk, c2 * 2 * mag[2] / hk]
        assert np.allclose(field, exp, rtol=0, atol=1e-14)


def test_field_shapes():
    # lib/field.cpp:37-54 closed forms: h sin(2 pi f t); +-h alternating each half period
    assert L.orc_field_sinusoidal(0.25, 2.0, 1.0) == pytest.approx(2.0, abs=1e-15)
    assert L.orc_field_sinusoidal(0.75, 2.0, 1.0) == pytest.approx(-2.0, abs=1e-15)
    assert L.orc_field_square(0.1, 3.0, 1.0) == 3.0
    assert L.orc_field_square(0.6, 3.0, 1.0) == -3.0
    assert L.orc_field_square(1.1, 3.0, 1.0) == 3.0


# --------------------------------------------------------------- distances
def test_pair_distance_vectors_known_answer():  # test/tests.cpp:906-946
    pts = f64([0, 0, 1, 3, 1, 0, 0.5, 0.5, 0.5])
    out = np.zeros(27)
    L.orc_pair_distance_vectors(ptr(out), ptr(pts), 3)
    d = out.reshape(3, 3, 3)
    assert np.all(d[[0, 1, 2], [0, 1, 2]] == 0)
    assert list(d[0, 1]) == [3, 1, -1] and list(d[1, 0]) == [-3, -1, 1]
    assert list(d[0, 2]) == [0.5, 0.5, -0.5] and list(d[2, 0]) == [-0.5, -0.5, 0.5]
    assert list(d[1, 2]) == [-2.5, -0.5, 0.5] and list(d[2, 1]) == [2.5, 0.5, -0.5]


def test_pair_distance_magnitudes_known_answer():  # test/tests.cpp:948-980
    d00, d01, d02, d12 = [0, 0, 0], [0, 0, 1], [3, 1, 0], [0.5, 0.5, 0.5]
    vecs = f64([d00, d01, d02, d01, d00, d12, d02, d12, d00]).reshape(-1)
    out = np.zeros(9)
    L.orc_pair_distance_magnitudes(ptr(out), ptr(vecs), 3)
    m = out.reshape(3, 3)
    assert m[0, 0] == m[1, 1] == m[2, 2] == 0
    assert m[0, 1] == 1 and m[1, 0] == 1
    assert dbl_eq(m[0, 2], 3.1622776601683795) and dbl_eq(m[2, 0], 3.1622776601683795)
    assert dbl_eq(m[1, 2], 0.8660254037844386) and dbl_eq(m[2, 1], 0.8660254037844386)


# --------------------------------------------------------------- trap
def test_trapezoidal_known_answer():  # test/tests.cpp:229-235
    x, y = f64([0.0, 0.1, 0.3, 0.7]), f64([5.0, 10.0, 10.0, 20.0])
    assert dbl_eq(L.orc_trapezoidal(ptr(x), ptr(y), 4), 8.75)


# --------------------------------------------------------------- RNG
def test_rng_mt_norm_known_answer():  # test/tests.cpp:346-355, RngMtNorm(999, 0.2)
    g = O.C.create_string_buffer(4096)
    L.orc_mtnorm_init.argtypes = [C.c_void_p, C.c_uint64, C.c_double]
    L.orc_mtnorm_get.restype = C.c_double
    L.orc_mtnorm_get.argtypes = [C.c_void_p]
    L.orc_mtnorm_init(g, 999, 0.2)
    got = [L.orc_mtnorm_get(g) for _ in range(5)]
    exp = [0.27100401702487437, -0.18950297511996847, -0.27620532277321042,
           -0.10214651921310165, -0.19125423381292103]
    for a, e in zip(got, exp):
        assert dbl_eq(a, e), (a, e)


def test_mt19937_64_standard_vector():
    # C++ [rand.predef]: the 10000th output of a default-seeded (5489)
    # mt19937_64 is 9981545732273789042.
    g = O.C.create_string_buffer(4096)
    L.orc_mtnorm_init.argtypes = [C.c_void_p, C.c_uint64, C.c_double]
    L.orc_mt64_next.restype = C.c_uint64
    L.orc_mt64_next.argtypes = [C.c_void_p]
    L.orc_mtnorm_init(g, 5489, 1.0)
    v = 0
    for _ in range(10000):
        v = L.orc_mt64_next(g)
    assert v == 9981545732273789042


@pytest.mark.parametrize("ctr,key,exp", [
    # Random123 kat_vectors, philox4x32 10 rounds
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox4x32_10_known_answer(ctr, key, exp):
    assert [int(v) for v in O.philox(ctr, key)] == exp


def test_wiener_stream_is_standard_normal():
    # Box-Muller over Philox: moments and a Kolmogorov-Smirnov check against N(0,1).
    from scipy import stats
    z = np.concatenate([O.wiener_normals(12345, p, 0, 20000) for p in range(5)])
    assert abs(z.mean()) < 4 / math.sqrt(z.size)
    assert abs(z.var() - 1) < 4 * math.sqrt(2 / z.size)
    assert stats.kstest(z, "norm").pvalue > 1e-3
    # distinct particles / tags / seeds give uncorrelated streams
    a = O.wiener_normals(1, 0, 0, 4000)
    b = O.wiener_normals(1, 1, 0, 4000)
    c = O.wiener_normals(2, 0, 0, 4000)
    d = O.wiener_normals(1, 0, 0, 4000, tag=1)
    for other in (b, c, d):
        assert abs(np.corrcoef(a, other)[0, 1]) < 0.07


def test_wiener_stream_random_access():
    # normal j depends only on (seed, particle, tag, j): any window agrees
    full = O.wiener_normals(77, 3, 0, 37)
    for j0 in (0, 1, 5, 17):
        assert np.array_equal(O.wiener_normals(77, 3, j0, 37 - j0), full[j0:])


# --------------------------------------------------------------- integrators
def _heun(sde_py, x0, w, n, wd, t, dt):
    nxt, dr, tdr = np.zeros(n), np.zeros(n), np.zeros(n)
    df, tdf = np.zeros(n * wd), np.zeros(n * wd)

    def cb(ctx, drift, diff, state, tt):
        s = np.ctypeslib.as_array(state, (n,))
        a, b = sde_py(s, tt)
        for i in range(n):
            drift[i] = a[i]
        for i in range(n * wd):
            diff[i] = b[i]

    fn = O.SDE_FN(cb)
    L.orc_heun(ptr(nxt), ptr(dr), ptr(tdr), ptr(df), ptr(tdf), ptr(f64(x0)), ptr(f64(w)), fn,
               None, n, wd, t, dt)
    return nxt


def test_heun_multiplicative_known_answer():  # test/tests.cpp:148-181
    def sde(x, t):
        return ([x[0] * x[1] * t, 3 * x[0]],
                [t, x[0], x[1], x[1], 4.0, x[0] * x[1]])

    dt = 0.1
    w = np.array([0.1, 0.01, 0.001]) / math.sqrt(dt)
    out = _heun(sde, [1, 2], w, 2, 3, 0.0, dt)
    assert dbl_eq(out[0], 1.03019352) and dbl_eq(out[1], 2.571186252)


def test_heun_driver_ou_pathwise():  # test/tests.cpp:183-227
    theta, mu, sigma, dt, n = 10.0, -1.0, 0.8, 1e-5, 5000
    g = O.C.create_string_buffer(4096)
    L.orc_mtnorm_init.argtypes = [C.c_void_p, C.c_uint64, C.c_double]
    L.orc_mtnorm_get.restype = C.c_double
    L.orc_mtnorm_get.argtypes = [C.c_void_p]
    L.orc_mtnorm_init(g, 1001, 1.0)
    w = [L.orc_mtnorm_get(g) for _ in range(n)]
    x = -3.0
    true = x
    for i in range(n):
        x = _heun(lambda s, t: ([theta * (mu - s[0])], [sigma]), [x], [w[i]], 1, 1,
                  i * dt, dt)[0]
        true = (true * math.exp(-theta * dt) + mu * (1 - math.exp(-theta * dt))
                + sigma * math.sqrt((1 - math.exp(-2 * theta * dt)) / (2 * theta)) * w[i])
        assert abs(true - x) < 1e-8


def _newton(fj_py, x0, eps, max_iter):
    n = len(x0)
    xr, xt, jac = np.zeros(n), np.zeros(n), np.zeros(n * n)
    ipiv = np.zeros(n, dtype=np.int32)
    err = C.c_int(0)

    def cb(ctx, fout, jout, xin):
        x = np.ctypeslib.as_array(xin, (n,)).copy()
        f, j = fj_py(x)
        for i in range(n):
            fout[i] = f[i]
        for i in range(n * n):
            jout[i] = j[i]

    fn = O.FJ_FN(cb)
    flag = L.orc_newton_noinv(ptr(xr), ptr(xt), ptr(jac), ptr(ipiv), C.byref(err), fn, None,
                              ptr(f64(x0)), n, eps, max_iter)
    return flag, xr, err.value


def _fj2(x):
    return ([-(x[0] - 2) ** 2 - 2 * x[0] * x[1], -x[1] * (x[0] + 1) - 4 * x[1]],
            [-2 * x[0] - 2 * x[1] + 4, -2 * x[0], -x[1], -x[0] - 5])


def test_newton_2d_two_solutions():  # test/tests.cpp:270-313
    eps = 1e-10
    flag, x, _ = _newton(_fj2, [-5.2, -5.2], eps, 100)
    tol = math.sqrt(2 * 5.2 * 5.2) * eps
    assert flag == 0 and abs(x[0] + 5) <= tol and abs(x[1] - 4.9) <= tol
    flag, x, _ = _newton(_fj2, [1.0, 1.0], eps, 100)
    tol = math.sqrt(2.0) * eps
    assert flag == 0 and abs(x[0] - 2) <= tol and abs(x[1]) <= tol


def test_newton_2d_singular():  # test/tests.cpp:315-344
    def fj(x):
        f0 = -(x[0] - 2) ** 2 - 2 * x[0] * x[1]
        j0, j1 = -2 * x[0] - 2 * x[1] + 4, -2 * x[0]
        return [f0, 2.0 * f0], [j0, j1, 2.0 * j0, 2.0 * j1]

    flag, _, err = _newton(fj, [-5.2, -5.2], 1e-8, 100)
    assert flag == 2 and err >= 0


def test_newton_max_iterations():
    # max_iter reached -> MAX_ITERATIONS_ERR (include/optimisation.hpp:22-33)
    flag, _, _ = _newton(_fj2, [-5.2, -5.2], 1e-14, 2)
    assert flag == 1


def test_dgesv_against_numpy_solve():
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 6, 9):
        for _ in range(20):
            A = rng.normal(size=(n, n))
            b = rng.normal(size=n)
            a, x = f64(A.copy()), f64(b.copy())
            ipiv = np.zeros(n, dtype=np.int32)
            assert L.orc_dgesv(n, ptr(a), ptr(ipiv), ptr(x)) == 0
            assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def _stiff_sde(a, b):
    def sde(x):
        return ([a * (x[1] - x[0]) - 0.5 * b * b * x[0], a * (x[0] - x[1]) - 0.5 * b * b * x[1]],
                [b * x[0], b * x[1]],
                [-a - 0.5 * b * b, a, a, -a - 0.5 * b * b],
                [b, 0, 0, b])
    return sde


def _implicit_step(sde_py, x0, dw, n, wd, t, dt, eps, max_iter):
    x, dwm, aw, bw = np.zeros(n), np.zeros(wd), np.zeros(n), np.zeros(n * wd)
    ad, bd = np.zeros(n * n), np.zeros(n * wd * n)
    xg, xt, xj = np.zeros(n), np.zeros(n), np.zeros(n * n)
    ipiv = np.zeros(n, dtype=np.int32)
    iters = C.c_int(0)

    def cb(ctx, a, b, adash, bdash, state, ta, tb):
        s = np.ctypeslib.as_array(state, (n,)).copy()
        A, B, AD, BD = sde_py(s)
        for i, v in enumerate(A):
            a[i] = v
        for i, v in enumerate(B):
            b[i] = v
        for i, v in enumerate(AD):
            adash[i] = v
        for i, v in enumerate(BD):
            bdash[i] = v

    fn = O.SDEJAC_FN(cb)
    code = L.orc_implicit_midpoint(ptr(x), ptr(dwm), ptr(aw), ptr(bw), ptr(ad), ptr(bd),
                                   ptr(xg), ptr(xt), ptr(xj), ptr(ipiv), ptr(f64(x0)),
                                   ptr(f64(dw)), fn, None, n, wd, t, dt, eps, max_iter,
                                   C.byref(iters))
    return code, x


def test_implicit_midpoint_stiff_2d():  # test/tests.cpp:434-481 (Kloeden & Platen p.397)
    code, x = _implicit_step(_stiff_sde(0.1, 5.0), [1.0, 2.0], [0.25], 2, 1, 0.0, 1e-4, 1e-6,
                             100)
    assert code == 0
    assert abs(x[0] - 1.01132363) < 5e-4 and abs(x[1] - 2.02261693) < 5e-4


def test_implicit_driver_stiff_2d_analytic():  # test/tests.cpp:483-539
    a, b, dt, n = 0.1, 5.0, 1e-4, 100
    g = O.C.create_string_buffer(4096)
    L.orc_mtnorm_init.argtypes = [C.c_void_p, C.c_uint64, C.c_double]
    L.orc_mtnorm_get.restype = C.c_double
    L.orc_mtnorm_get.argtypes = [C.c_void_p]
    L.orc_mtnorm_init(g, 1001, 1.0)
    dw = [L.orc_mtnorm_get(g) for _ in range(n)]
    wt = np.concatenate([[0.0], np.cumsum(dw[:-1])])
    x = np.array([1.0, 2.0])
    x0 = x.copy()
    sde = _stiff_sde(a, b)
    t = 0.0
    for i in range(1, n):
        t += dt
        code, x = _implicit_step(sde, x, [dw[i - 1]], 2, 1, t, dt, 1e-6, 500)
        rp = -0.5 * b * b * i * dt + b * wt[i] * math.sqrt(dt)
        rm = (-2 * a - 0.5 * b * b) * i * dt + b * wt[i] * math.sqrt(dt)
        on, off = math.exp(rp) + math.exp(rm), math.exp(rp) - math.exp(rm)
        assert abs(0.5 * (on * x0[0] + off * x0[1]) - x[0]) < 0.002
        assert abs(0.5 * (off * x0[0] + on * x0[1]) - x[1]) < 0.002
