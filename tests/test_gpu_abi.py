"""mp_full_dynamics, the drop-in for simulation::full_dynamics
(include/simulation.hpp:74-94), called through ctypes exactly as the
reference side would bind it (no build step, plain pointers and sizes).

A two-particle interacting cluster in FP64 under a sine field: the result
equals core.simulate's (the Cython path into the same library) bit for bit
and the oracle's full_dynamics on the same Philox increments to 1e-10 of Ms.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

core = pytest.importorskip("magpy_amd.core")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MU0 = 1.25663706e-6


def bind():
    lib = C.CDLL(os.path.join(ROOT, "magpy_amd", "libmagpy_amd.so"))
    dp = C.POINTER(C.c_double)
    lib.mp_full_dynamics.argtypes = [dp, dp, dp, dp, dp, C.c_size_t, C.c_double, C.c_double,
                                     C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                                     C.c_double, C.c_double, C.c_size_t, C.c_long, C.c_int,
                                     C.c_double, C.c_double, C.c_int, dp, dp, dp, dp, dp]
    lib.mp_full_dynamics.restype = C.c_int
    lib.mp_last_error.restype = C.c_char_p
    return lib, (lambda a: a.ctypes.data_as(dp))


def full_dynamics(lib, P, r, K, axes, m0, loc, implicit, S, end, dt, seed, precision=1):
    n = len(r)
    t, f = np.zeros(S), np.zeros(S)
    x, y, z = (np.zeros((n, S)) for _ in range(3))
    rc = lib.mp_full_dynamics(P(r), P(K), P(axes), P(m0), P(loc), n, 4e5, 0.1, 300.0, 0, 1,
                              int(implicit), 1e-9, dt, end, S, seed, 0, 0.02 / MU0, 3e9,
                              precision, P(t), P(f), P(x), P(y), P(z))
    return rc, dict(time=t, field=f, x=x, y=y, z=z)


def dimer():
    r = np.array([10e-9, 11e-9])
    K = np.array([23e3, 20e3])
    axes = np.array([[0.0, 0.6, 0.8], [0.8, 0.0, 0.6]])
    m0 = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8]])
    loc = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 22e-9]])
    return r, K, axes, m0, loc


def test_full_dynamics_ctypes_matches_binding_and_oracle():
    print("CASE abi_full_dynamics_heun", flush=True)
    lib, P = bind()
    r, K, axes, m0, loc = dimer()
    dt, S, seed = 1e-14, 21, 1001
    end = 2000 * dt
    rc, g = full_dynamics(lib, P, r, K, axes, m0, loc, False, S, end, dt, seed)
    assert rc == 0, lib.mp_last_error()
    c = core.simulate(r, K, axes, m0, loc, 4e5, 0.1, 300.0, False, True, False, dt, end, S, seed,
                      'sine', 0.02 / MU0, 3e9, 1e-9, precision='fp64')
    np.testing.assert_array_equal(g["time"], c["time"])
    np.testing.assert_array_equal(g["field"], c["field"])
    for comp in "xyz":
        for i in range(2):
            np.testing.assert_array_equal(g[comp][i], c[comp][i])
    o = O.full_dynamics(r, K, axes, m0, loc, 4e5, 0.1, 300.0, False, True, False, 1e-9, dt, end,
                        S, seed, field_shape='sine', field_amplitude=0.02 / MU0,
                        field_frequency=3e9, rng='philox')
    np.testing.assert_array_equal(g["time"], o["time"])
    np.testing.assert_allclose(g["field"], o["field"], rtol=1e-15, atol=0)
    for comp in "xyz":
        np.testing.assert_allclose(g[comp], o[comp], rtol=0, atol=1e-10 * 4e5)
    # the dimer interacts: the run without the dipolar term differs
    o_free = O.full_dynamics(r, K, axes, m0, loc, 4e5, 0.1, 300.0, False, False, False, 1e-9,
                             dt, end, S, seed, field_shape='sine', field_amplitude=0.02 / MU0,
                             field_frequency=3e9, rng='philox')
    assert np.max(np.abs(o_free["z"] - g["z"])) > 1e-6 * 4e5


def test_full_dynamics_ctypes_implicit_dimer():
    # the reference's default path: implicit midpoint with interactions
    print("CASE abi_full_dynamics_implicit", flush=True)
    lib, P = bind()
    r, K, axes, m0, loc = dimer()
    dt, S, seed = 5e-13, 11, 77
    end = 300 * dt
    rc, g = full_dynamics(lib, P, r, K, axes, m0, loc, True, S, end, dt, seed)
    assert rc == 0, lib.mp_last_error()
    o = O.full_dynamics(r, K, axes, m0, loc, 4e5, 0.1, 300.0, False, True, True, 1e-9, dt, end,
                        S, seed, field_shape='sine', field_amplitude=0.02 / MU0,
                        field_frequency=3e9, rng='philox')
    for comp in "xyz":
        np.testing.assert_allclose(g[comp], o[comp], rtol=0, atol=1e-10 * 4e5)


def test_full_dynamics_ctypes_errors():
    print("CASE abi_full_dynamics_errors", flush=True)
    lib, P = bind()
    r, K, axes, m0, loc = dimer()
    rc, _ = full_dynamics(lib, P, r, K, axes, m0, loc, False, 0, 1e-12, 1e-14, 1)
    assert rc == 1 and b"max_samples" in lib.mp_last_error()     # MP_ERR_ARG
    rc, _ = full_dynamics(lib, P, r, K, axes, m0, loc, False, 5, 1e-12, 1e-14, 1, precision=7)
    assert rc == 1 and b"precision" in lib.mp_last_error()
