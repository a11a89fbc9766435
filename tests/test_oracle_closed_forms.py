"""The oracle's full_dynamics orchestration against closed-form solutions.

owlas/magpy's simulation::full_dynamics (lib/simulation.cpp:476-624 and
:132-429) cannot be compiled here (it includes cblas.h / lapacke.h), so its
orchestration is pinned by zero-temperature runs whose answers follow from
the mathematics alone:

- the unit reduction and time factor gamma mu0 H_k / (1 + alpha^2)
  (lib/simulation.cpp:527-533): uniaxial relaxation with the easy axis along
  z obeys d(m_z^2)/dt' = 2 alpha m_z^2 (1 - m_z^2) in reduced time t', a
  logistic equation, and the azimuth advances by the integral of m_z;
- the field's time convention (step n -> n+1 evaluates the field at
  t_{n+1} and t_{n+2}, lib/simulation.cpp:350-375): with alpha = 0 and all
  fields along z the Heun step is a rotation about z whose complex
  amplification 1 + i dt (w1 + w2) / 2 - dt^2 w1 w2 / 2 is exact for the
  scheme, so the discrete trajectory is a closed-form product;
- the zero-order-hold sample index (lib/simulation.cpp:342-405): sample s is
  the state after floor(s Ts / dt) steps, and the run takes one step more
  than the last sample's;
- the reduced -> SI conversion (lib/simulation.cpp:610-621): time / factor,
  field * H_k, m * Ms.

Each check uses inputs in SI units, so an error in the conversion (a dropped
1/(1+alpha^2), the field taken at t_n, the state one step late) moves the
result by orders of magnitude more than the stated tolerance.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

MU0, GAMMA = 1.25663706e-6, 1.76086e11   # include/constants.hpp:11-12
K, MS, RADIUS = 23e3, 4e5, 10e-9
HK = 2 * K / (MU0 * MS)                 # anisotropy field, A/m


def time_factor(alpha):
    return GAMMA * MU0 * HK / (1 + alpha * alpha)


def run(alpha, m0, dt_red, end_red, samples, field_shape="constant", h_red=0.0, f_red=0.0):
    """One particle, easy axis along z, T = 0, from reduced targets given in SI."""
    tf = time_factor(alpha)
    dt, end = dt_red / tf, end_red / tf
    o = O.full_dynamics([RADIUS], [K], [[0, 0, 1]], [m0], [[0, 0, 0]], MS, alpha, 0.0,
                        False, False, False, 1e-9, dt, end, samples, 7, field_shape=field_shape,
                        field_amplitude=h_red * HK, field_frequency=f_red * tf, rng="philox")
    return o, dt, end


def zoh_steps(dt, end, samples):
    """k_s = floor(s Ts / dt): the number of steps taken before sample s."""
    ts = end / (samples - 1)
    return np.array([math.floor(s * ts / dt) for s in range(samples)], dtype=np.int64)


# ------------------------------------------------------- uniaxial relaxation
ALPHA, THETA0 = 0.5, 1.2
SAMPLES = 41


def relaxation(t_red):
    """m_z(t') and the azimuth advance for m_z(0) = cos(THETA0), axis = z."""
    c = 1.0 / math.cos(THETA0) ** 2 - 1.0
    mz = 1.0 / np.sqrt(1.0 + c * np.exp(-2 * ALPHA * t_red))
    phi = (np.arcsinh(np.exp(ALPHA * t_red) / math.sqrt(c)) - math.asinh(1 / math.sqrt(c))) / ALPHA
    return mz, phi


@pytest.mark.parametrize("dt_red", [4e-4, 2e-4])
def test_uniaxial_relaxation_time_factor_and_sample_index(dt_red):
    end_red = 7.9937   # Ts / dt is not an integer: the floor in k_s matters
    o, dt, end = run(ALPHA, [math.sin(THETA0), 0.0, math.cos(THETA0)], dt_red, end_red, SAMPLES)
    k = zoh_steps(dt, end, SAMPLES)
    assert o["stats"].n_steps == k[-1] + 1
    t_red = k * dt * time_factor(ALPHA)
    mz_exact, phi_exact = relaxation(t_red)
    mz = o["z"][0] / MS
    phi = np.unwrap(np.arctan2(o["y"][0], o["x"][0]))
    tol = 0.1 * dt_red ** 2           # Heun's global error is 0.07 dt^2 here
    assert np.max(np.abs(mz - mz_exact)) < tol
    assert np.max(np.abs(phi - phi_exact)) < 10 * tol
    # |m| stays 1 to the scheme's order (no renormalisation asked for)
    nrm = np.sqrt(o["x"][0] ** 2 + o["y"][0] ** 2 + o["z"][0] ** 2) / MS
    assert np.max(np.abs(nrm - 1)) < tol
    # the state one step later is far outside the tolerance: the index is pinned
    mz_late, _ = relaxation((k + 1) * dt * time_factor(ALPHA))
    assert np.max(np.abs(mz - mz_late)) > 100 * tol
    # SI time axis: s * Ts in seconds
    np.testing.assert_allclose(o["time"], np.arange(SAMPLES) * (end / (SAMPLES - 1)),
                               rtol=1e-13, atol=0)


def test_relaxation_is_second_order_in_dt():
    """Halving dt quarters the distance to the closed form (Heun is a
    second-order scheme for the deterministic equation)."""
    errs = []
    for dt_red in (4e-4, 2e-4, 1e-4):
        o, dt, end = run(ALPHA, [math.sin(THETA0), 0.0, math.cos(THETA0)], dt_red, 7.9937,
                         SAMPLES)
        k = zoh_steps(dt, end, SAMPLES)
        mz_exact, _ = relaxation(k * dt * time_factor(ALPHA))
        errs.append(np.max(np.abs(o["z"][0] / MS - mz_exact)))
    assert 3.6 < errs[0] / errs[1] < 4.4 and 3.6 < errs[1] / errs[2] < 4.4


# ---------------------------------------- precession under a sinusoidal field
H_RED, F_RED, TH_P = 0.8, 1 / 1.3, 0.9
DT_P, END_P, S_P = 1e-3, 6.4979, 51


def heun_rotation(n_steps, dt, mz, offset):
    """The scheme's exact trajectory for alpha = 0, fields along z: rotation
    rate w(t) = mz + h sin(2 pi f t), one Heun step n -> n+1 multiplies
    m_x + i m_y by 1 + i dt (w1 + w2)/2 - dt^2 w1 w2 / 2 with the rates taken
    at the times t_{n+offset} and t_{n+offset} + dt."""
    n = np.arange(n_steps, dtype=np.float64)
    t1 = (n + offset) * dt
    w1 = mz + H_RED * np.sin(2 * np.pi * F_RED * t1)
    w2 = mz + H_RED * np.sin(2 * np.pi * F_RED * (t1 + dt))
    amp = 1 + 1j * dt * (w1 + w2) / 2 - dt * dt * w1 * w2 / 2
    z = np.concatenate([[1.0 + 0j], np.cumprod(amp)])
    return z * math.sin(TH_P)


def test_precession_field_time_convention():
    o, dt, end = run(0.0, [math.sin(TH_P), 0.0, math.cos(TH_P)], DT_P, END_P, S_P,
                     field_shape="sine", h_red=H_RED, f_red=F_RED)
    k = zoh_steps(dt, end, S_P)
    assert o["stats"].n_steps == k[-1] + 1
    z = (o["x"][0] + 1j * o["y"][0]) / MS
    mz = math.cos(TH_P)
    np.testing.assert_array_equal(o["z"][0] / MS, np.full(S_P, o["z"][0][0] / MS))
    ref = heun_rotation(k[-1] + 1, DT_P, mz, 1)[k]       # field at t_{n+1}, t_{n+2}
    assert np.max(np.abs(z - ref)) < 1e-12
    early = heun_rotation(k[-1] + 1, DT_P, mz, 0)[k]     # field at t_n, t_{n+1}
    assert np.max(np.abs(z - early)) > 1e-4
    # the continuous solution with the field shifted by one step, to O(dt^2);
    # without the shift it is O(dt) away
    t = k * DT_P
    w = 2 * np.pi * F_RED

    def phase(shift):
        return mz * t + H_RED * (np.cos(w * shift) - np.cos(w * (t + shift))) / w

    dphi = np.unwrap(np.angle(z))
    assert np.max(np.abs(dphi - phase(DT_P))) < 5e-6
    assert np.max(np.abs(dphi - phase(0.0))) > 3e-4
    # SI applied field at the nominal sample times
    ts = end / (S_P - 1)
    np.testing.assert_allclose(o["field"], H_RED * HK * np.sin(2 * np.pi * F_RED
                                                               * time_factor(0.0)
                                                               * np.arange(S_P) * ts),
                               rtol=0, atol=1e-9 * HK)


# -------------------------------------------------------------------- dnrm2
def _ulp_close(a, b):
    return abs(a - b) <= np.spacing(abs(b))


@pytest.mark.parametrize("scale", [1.0, 1e-3, 1e150, 1e300, 1e-150, 1e-300, 1e-310])
def test_dnrm2_within_one_ulp(scale):
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 5, 9):
        for _ in range(200):
            x = rng.normal(size=n) * scale
            assert _ulp_close(O.dnrm2(x), math.hypot(*x)), (x, O.dnrm2(x), math.hypot(*x))


def test_dnrm2_special_values():
    assert O.dnrm2(np.zeros(3)) == 0.0
    assert O.dnrm2(np.array([3.0, 4.0, 12.0])) == 13.0
    big = np.finfo(np.float64).max
    assert O.dnrm2(np.array([big, 0.0, 0.0])) == big
    tiny = 5e-324
    assert O.dnrm2(np.array([tiny, 0.0, 0.0])) == tiny
