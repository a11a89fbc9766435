"""Physics of the device integrator against what the mathematics fixes.

- Equilibrium: a particle in its easy-axis well (no field) relaxes to the
  Boltzmann density p(m_z) ~ exp(sigma m_z^2) on [0, 1]; its CDF and mean
  have closed forms through erfi (Z = sqrt(pi / 4 sigma) erfi(sqrt(sigma)),
  E[m_z] = (e^sigma - 1) / (2 sigma Z)).  The config-1 particle (r = 12 nm,
  sigma = 40.19) is run past its intra-well relaxation and its histogram is
  compared with that density (chi-square over equiprobable bins).
- Weak error of the equilibrium: the bias of <1 - m_z> vanishes as dt -> 0
  (within the sampling error at dt_red = 0.0125) and grows at least linearly
  with dt.
- Strong order: on one Brownian path per particle, coarse-grained by summing
  fine increments (w_coarse = sum w_fine / sqrt(m), the construction of
  test/convergence/task5.cpp:108-121), the mean pathwise error against the
  finest level falls with order >= 0.45 (Heun's strong order for this
  multiplicative noise is 1/2; the reference reports 0.509 and 0.589 for its
  task).
- Deterministic order: with T = 0 the scheme is Heun's second-order method
  (the deterministic task of test/convergence/task4.cpp).
"""
import numpy as np
import pytest
from scipy.special import erfi
from scipy.stats import chi2

pytestmark = pytest.mark.gpu

core = pytest.importorskip("magpy_amd.core")

MU0 = 1.25663706e-6
R, K, MS, AL = 12e-9, 23e3, 4e5, 0.1
TAU = 1.76086e11 * MU0 * (2 * K / MU0 / MS) / (1 + AL * AL)   # reduced time per second
SIGMA = K * 4 / 3 * np.pi * R ** 3 / (1.38064852e-23 * 300.0)


def ensemble(P, m0, h_red, T, seed=1):
    rng = np.random.default_rng(seed)
    return dict(member_offsets=np.arange(P + 1, dtype=np.int64), radius=np.full(P, R),
                anisotropy=np.full(P, K), anisotropy_axis=np.tile([0, 0, 1.0], (P, 1)),
                magnetisation_direction=np.tile(m0, (P, 1)), location=np.zeros((P, 3)),
                magnetisation=MS, damping=AL, temperature=T, field_shape="constant",
                field_amplitude=h_red * 2 * K / MU0 / MS, field_frequency=0.0,
                seeds=rng.integers(1, 2 ** 31 - 1, size=P))


def boltzmann_cdf(x):
    return erfi(np.sqrt(SIGMA) * x) / erfi(np.sqrt(SIGMA))


def boltzmann_mean_gap():
    Z = np.sqrt(np.pi / (4 * SIGMA)) * erfi(np.sqrt(SIGMA))
    return 1.0 - (np.exp(SIGMA) - 1) / (2 * SIGMA * Z)


def last_sample_mz(o):
    """Direction m_z of every particle at the last sample."""
    v = np.stack([o["x"][:, -1], o["y"][:, -1], o["z"][:, -1]], 1)
    return v[:, 2] / np.linalg.norm(v, axis=1)


def test_config1_equilibrium_matches_boltzmann():
    print("CASE boltzmann_config1", flush=True)
    P = 65536
    e = ensemble(P, [0, 0, 1.0], 0.0, 300.0, seed=5)
    # 4 ns: the intra-well variance relaxes at 2 alpha / (1 + alpha^2) per
    # reduced time unit (~0.25 ns), so 4 ns leaves e^-16 of the start
    o = core.simulate_ensemble(time_step=1e-14, end_time=4e-9, max_samples=3, precision="fp32",
                               **e)
    mz = last_sample_mz(o)
    gap, se = np.mean(1 - mz), np.std(mz) / np.sqrt(P)
    print("mean 1-mz", gap, "exact", boltzmann_mean_gap(), "se", se, flush=True)
    assert abs(gap - boltzmann_mean_gap()) < 4 * se
    edges = np.interp(np.linspace(0, 1, 21), boltzmann_cdf(np.linspace(0, 1, 200001)),
                      np.linspace(0, 1, 200001))
    edges[0], edges[-1] = -np.inf, np.inf
    counts, _ = np.histogram(mz, bins=edges)
    stat = np.sum((counts - P / 20) ** 2 / (P / 20))
    p = chi2.sf(stat, 19)
    print("chi2", stat, "p", p, flush=True)
    assert p > 1e-3


def test_equilibrium_bias_is_a_time_step_error():
    print("CASE boltzmann_weak_order", flush=True)
    P = 1 << 22
    e = ensemble(P, [0, 0, 1.0], 0.0, 300.0, seed=6)
    bias, se = {}, {}
    for dtr in (0.0125, 0.1, 0.2):
        n = int(round(120 / dtr))
        o = core.simulate_ensemble(time_step=dtr / TAU, end_time=(n + 0.5) * dtr / TAU,
                                   max_samples=2, precision="fp32", trajectories=True, **e)
        mz = last_sample_mz(o)
        bias[dtr] = np.mean(1 - mz) - boltzmann_mean_gap()
        se[dtr] = np.std(mz) / np.sqrt(P)
    print("bias", bias, "se", se, flush=True)
    assert abs(bias[0.0125]) < 4 * se[0.0125]
    assert abs(bias[0.2]) > 8 * se[0.2] and abs(bias[0.1]) > 3 * se[0.1]
    assert np.log2(bias[0.2] / bias[0.1]) > 0.8


def levels(e, dt_fine, n_fine, mults, precision="fp64", noise=None):
    out = {}
    P = len(e["radius"])
    for m in mults:
        n = n_fine // m
        kw = {}
        if noise is not None:
            w = noise.reshape(n, m, P, 3).sum(1) / np.sqrt(m)
            kw["noise"] = np.concatenate([w, np.zeros((1, P, 3))])  # the loop's extra step
        o = core.simulate_ensemble(time_step=dt_fine * m, end_time=(n + 0.5) * dt_fine * m,
                                   max_samples=2, precision=precision, **kw, **e)
        out[m] = np.stack([o["x"][:, -1], o["y"][:, -1], o["z"][:, -1]], 1) / MS
    return out


def test_strong_order_on_coarse_grained_paths():
    print("CASE strong_order", flush=True)
    P, nf, dtr = 4096, 6400, 1e-3
    e = ensemble(P, [1.0, 0, 0], 0.1, 300.0)
    noise = np.random.default_rng(7).normal(size=(nf, P, 3))
    res = levels(e, dtr / TAU, nf, (1, 2, 4, 8, 16, 32, 64), noise=noise)
    ms = [4, 8, 16, 32, 64]
    err = [np.mean(np.linalg.norm(res[m] - res[1], axis=1)) for m in ms]
    slope = np.polyfit(np.log(ms), np.log(err), 1)[0]
    print("strong errors", dict(zip(ms, err)), "order", slope, flush=True)
    assert np.all(np.diff(err) > 0)
    assert 0.45 < slope < 1.0


def test_deterministic_order_two():
    print("CASE deterministic_order", flush=True)
    P = 64
    e = ensemble(P, [1.0, 0, 0], 0.3, 0.0)
    m0 = np.random.default_rng(3).normal(size=(P, 3))
    e["magnetisation_direction"] = m0 / np.linalg.norm(m0, axis=1, keepdims=True)
    res = levels(e, 1e-3 / TAU, 6400, (1, 2, 4, 8, 16, 32, 64))
    ms = [4, 8, 16, 32, 64]
    err = [np.max(np.linalg.norm(res[m] - res[1], axis=1)) for m in ms]
    slope = np.polyfit(np.log(ms), np.log(err), 1)[0]
    print("deterministic errors", dict(zip(ms, err)), "order", slope, flush=True)
    assert 1.8 < slope < 2.3


def test_config1_plain_fp64_ensemble():
    """The reference's CPU-runnable case: one 12 nm particle (K = 23 kJ/m^3,
    Ms = 400 kA/m, alpha = 0.1, 300 K, axis and m0 along +z, no field), 1,000
    realisations with their own seeds, 1e5 Heun steps of 1e-14 s in FP64.
    Three realisations equal the oracle run on the same Philox stream over
    the whole time series; the 1,000 final m_z are tested against the
    Boltzmann distribution on [0, 1] (Kolmogorov-Smirnov)."""
    print("CASE config1_plain_fp64", flush=True)
    from scipy.stats import kstest
    from oracle import oracle as O
    M, S = 1000, 101
    e = ensemble(M, [0, 0, 1.0], 0.0, 300.0, seed=9)
    dt, end = 1e-14, 1e-9
    o = core.simulate_ensemble(time_step=dt, end_time=end, max_samples=S, precision="fp64",
                               trajectories=True, **e)
    assert np.all(o["status"] == 0)
    for m in (0, 499, 999):
        r = O.full_dynamics([R], [K], [[0, 0, 1.0]], [[0, 0, 1.0]], [[0, 0, 0]], MS, AL, 300.0,
                            False, False, False, 1e-9, dt, end, S, int(e["seeds"][m]),
                            rng="philox")
        for c in "xyz":
            assert np.allclose(o[c][m], r[c][0], rtol=0, atol=MS * 1e-10), (m, c)
    mz = last_sample_mz(o)
    p = kstest(mz, boltzmann_cdf).pvalue
    print("config 1: mean 1-mz", np.mean(1 - mz), "exact", boltzmann_mean_gap(), "KS p", p,
          flush=True)
    assert p > 1e-3
