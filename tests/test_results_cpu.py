"""Results / EnsembleResults and the SAR estimator against closed forms.

The hysteresis-loop energy of owlas/magpy (magpy/results.py:167-217) is
E = -mu0 * trapz(y=H, x=<M>) over a time window, the final cycle being
t >= t_end - 1/f.  For H = H0 sin(wt) and <M> = M0 sin(wt - d) sampled n
times per period, the trapezoid sum over one whole period is exactly
pi H0 M0 sin(d) sin(h)/h with h = 2 pi / n (the end terms telescope and each
interval contributes H0 M0 sin(d) sin(h) / 2), so E has a closed form that a
swapped integration variable (sign), a missing mu0, a wrong window or an
unweighted-vs-weighted mean would all miss by far more than rounding.
"""
import math

import numpy as np
import pytest

from magpy_amd import core
from magpy_amd.results import EnsembleResults, Results, trapezoid

MU0 = 1.25663706e-6   # include/constants.hpp:11


def test_trapezoid_known_answer():   # test/tests.cpp:229-235
    assert trapezoid([5.0, 10.0, 10.0, 20.0], [0.0, 0.1, 0.3, 0.7]) == pytest.approx(8.75,
                                                                                      rel=1e-15)


def test_mu0():
    assert core.get_mu0() == MU0


DT_SAMPLE = 2.0 ** -21   # s; sample times k * DT_SAMPLE and t_end - 1/f are exact


def loop_ensemble(n_per_cycle, cycles, h0, m0, delta):
    """Three members of two particles whose mean total z magnetisation is
    m0 sin(wt - delta); member-to-member differences cancel in the mean.
    The period is n_per_cycle samples, so the final cycle's window starts
    exactly on a sample (the reference keeps t >= t_end - 1/f)."""
    n = n_per_cycle * cycles + 1
    freq = 1.0 / (n_per_cycle * DT_SAMPLE)
    assert 1.0 / freq == n_per_cycle * DT_SAMPLE
    t = np.arange(n) * DT_SAMPLE
    th = 2 * np.pi * freq * t
    field = h0 * np.sin(th)
    target = m0 * np.sin(th - delta)
    eps = 0.3 * m0 * np.cos(3 * th + 0.2)
    members = []
    for share, wobble in ((0.25, eps), (0.5, -eps), (0.25, 0.0)):
        z0 = share * target + wobble
        z1 = (1 - share) * target
        members.append(Results(time=t, field=field,
                               x={0: np.cos(th), 1: -np.cos(th)}, y={0: np.sin(th), 1: 0 * th},
                               z={0: z0, 1: z1}, N=2))
    return EnsembleResults(members), t, freq


@pytest.mark.parametrize("n_per_cycle,delta", [(64, 0.3), (200, 1.1), (37, -0.4)])
def test_final_cycle_energy_closed_form(n_per_cycle, delta):
    h0, m0 = 15915.5, 2.4e5
    ens, t, freq = loop_ensemble(n_per_cycle, 3, h0, m0, delta)
    np.testing.assert_allclose(ens.ensemble_magnetisation('z'),
                               m0 * np.sin(2 * np.pi * freq * t - delta), rtol=0,
                               atol=1e-9 * m0)
    h = 2 * np.pi / n_per_cycle
    per_cycle = -MU0 * math.pi * h0 * m0 * math.sin(delta) * math.sin(h) / h
    e_final = ens.final_cycle_energy_dissipated(freq)
    assert e_final == pytest.approx(per_cycle, rel=1e-9)
    # the sampled loop is not the continuous one: sin(h)/h is resolved
    assert abs(e_final / (-MU0 * math.pi * h0 * m0 * math.sin(delta)) - 1) > 0.1 * h * h / 6
    # whole run: three cycles; a window of one interior cycle
    assert ens.energy_dissipated() == pytest.approx(3 * per_cycle, rel=1e-9)
    T = 1 / freq
    assert ens.energy_dissipated(start_time=t[n_per_cycle], end_time=t[2 * n_per_cycle]) == \
        pytest.approx(per_cycle, rel=1e-9)
    # a quarter cycle is not a closed loop: its end terms H M / 2 remain
    part = ens.energy_dissipated(start_time=t[-1] - 0.25 * T)
    assert abs(part - per_cycle / 4) > 0.02 * abs(per_cycle)


def test_ensemble_magnetisation_is_unweighted_mean_of_sums():
    """<M> is the plain mean over members of the sum over particles
    (magpy/results.py:116-151), whatever the particle sizes."""
    t = np.linspace(0, 1, 5)
    mk = lambda zs: Results(t, 0 * t, {i: 0 * t for i in range(len(zs))},  # noqa: E731
                            {i: 0 * t for i in range(len(zs))},
                            {i: np.full(5, z) for i, z in enumerate(zs)}, len(zs))
    ens = EnsembleResults([mk([1.0, 2.0]), mk([4.0]), mk([0.5, 0.5, 0.5])])
    np.testing.assert_array_equal(ens.ensemble_magnetisation('z'), np.full(5, (3 + 4 + 1.5) / 3))
    assert [float(m[0]) for m in ens.magnetisation('z')] == [3.0, 4.0, 1.5]
    assert ens.final_state()[0]['z'] == {0: 1.0, 1: 2.0}


def test_plot_labels_time_in_seconds():
    matplotlib = pytest.importorskip("matplotlib")
    matplotlib.use("Agg")
    t = np.linspace(0, 1e-9, 4)
    r = Results(t, 0 * t, {0: t, 1: t}, {0: t, 1: t}, {0: t, 1: t}, 2)
    fig = r.plot()
    axes = fig.get_axes()
    assert len(axes) == 2 and axes[1].get_xlabel() == 'Time [s]'
    assert [ln.get_label() for ln in axes[0].get_lines()] == ['x', 'y', 'z']
