"""The oracle's discrete orientation model pinned by the reference's own
known answers (test/tests.cpp:544-616: rk45 steps, the master equation, the
Neel-Brown transition matrix) and by the closed form of a two-state master
equation with constant rates."""
import numpy as np

from oracle import oracle as O


def ode(x, t):
    return np.array([x[1] * t, -x[0] - 5])


def test_rk45_time_dependent_step_known_answer():
    # test/tests.cpp:566-593 (the same ODE at t = 0.5, dt = 0.01, eps = 1e-4)
    x2, ks, t, h = O.rk45(ode, [1.0, 2.0], 0.5, 0.01, 1e-4)
    k1, k2, k3, k4, k5, k6 = ks
    expect = [(1.0, k1[0]), (-6.0, k1[1]), (0.997976, k2[0]), (-6.002, k2[1]),
              (0.9969437365, k3[0]), (-6.002995446, k3[1]), (0.993774919651888, k4[0]),
              (-6.005981540838, k4[1]), (0.9893745618215710, k5[0]),
              (-6.009947940975117, k5[1]), (0.99077120433373344621, k6[0]),
              (-6.00871032589945830438, k6[1]), (1.00994791563354557873, x2[0]),
              (1.93995017187702467609, x2[1])]
    for want, got in expect:
        # ASSERT_DOUBLE_EQ: within 4 ulp
        assert abs(got - want) <= 4 * np.spacing(abs(want)), (want, got)


def test_master_equation_known_answer():
    # test/tests.cpp:595-605
    W = np.arange(1.0, 10.0)
    out = np.zeros(3)
    O.lib().orc_master_equation(O.ptr(out), O.ptr(W), O.ptr(np.array([3.0, 2.0, 1.0])), 3)
    assert np.array_equal(out, [10.0, 28.0, 46.0])


def test_transition_matrix_known_answer():
    # test/tests.cpp:607-616
    W = np.zeros(4)
    O.lib().orc_dom_transition_matrix(O.ptr(W), 1.0, 2e-21, 10.0, 0.5, 2.7, 0.1)
    want = [-0.00021827108536028533, 278104386.1896516, 0.00021827108536028533,
            -278104386.1896516]
    for w, g in zip(want, W):
        assert abs(g - w) <= 4 * np.spacing(abs(w)), (w, g)


def test_constant_field_matches_closed_form():
    """dp0/dt = -r2 p0 + r1 p1, p0 + p1 = 1: mz(t) = mz_eq + (mz0 - mz_eq) e^{-(r1+r2) t}
    with mz_eq = (r1 - r2) / (r1 + r2)."""
    K, r, T, ms, al, h = 23e3, 7e-9, 300.0, 4e5, 0.1, 0.2
    V = 4 / 3 * np.pi * r ** 3
    W = np.zeros(4)
    O.lib().orc_dom_transition_matrix(O.ptr(W), K, V, T, h, ms, al)
    r2, r1 = -W[0], W[1]
    lam = r1 + r2
    end = 5.0 / lam
    res = O.dom_ensemble_dynamics(V, K, T, ms, al, "constant", h, 0.0, [1.0, 0.0], 1e-12, end, 201)
    mz_eq = (r1 - r2) / lam
    exact = mz_eq + (1.0 - mz_eq) * np.exp(-lam * res["time"])
    assert np.max(np.abs(res["mz"] - exact)) < 1e-6
    assert np.all(res["field"] == h)


def test_square_fourier_matches_its_series():
    f, h = 3e5, 0.3
    t = np.linspace(0, 1 / f, 7)
    for n in (1, 3):
        got = [O.lib().orc_field_square_fourier(x, h, f, n) for x in t]
        series = 4 / np.pi * h * sum(np.sin(2 * np.pi * (2 * k - 1) * f * t) / (2 * k - 1)
                                     for k in range(1, n + 1))
        assert np.allclose(got, series, rtol=0, atol=1e-15)
        ref = O.ref()
        if ref is not None:   # the reference's own field::square_fourier, bit for bit
            assert got == [ref.ref_field_square_fourier(x, h, f, n) for x in t]
