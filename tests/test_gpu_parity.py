"""Parity of the HIP path (through the C ABI / Cython binding) with the FP64
CPU oracle on the same inputs and the same Wiener increments.

Tolerances (stated per test): FP64 device vs FP64 oracle differ only in
operation order (the device folds drift and diffusion into one bilinear map,
uses FMA, and evaluates transcendentals with the device math library), so
trajectories agree to ~1e-12 relative over thousands of steps; FP32 device
runs are compared with the FP64 device run on the same Philox stream.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

core = pytest.importorskip("magpy_amd.core")

KB, MU0 = 1.38064852e-23, 1.25663706e-6


def _print_case(name, **kw):
    print("CASE", name, kw, flush=True)


def ensemble_inputs(M, seed=0, n_per=1, field_shape="sine", radius=None, K=None):
    rng = np.random.default_rng(seed)
    P = M * n_per
    offs = np.arange(M + 1, dtype=np.int64) * n_per
    r = radius if radius is not None else 10e-9 * np.exp(0.1 * rng.normal(size=P))
    K = np.full(P, 23e3) if K is None else K
    th = 2 * np.pi * rng.uniform(size=P)
    ph = np.arccos(1 - 2 * rng.uniform(size=P))
    axes = np.stack([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)], 1)
    return dict(member_offsets=offs, radius=np.asarray(r, float), anisotropy=np.asarray(K, float),
                anisotropy_axis=axes, magnetisation_direction=axes.copy(),
                location=np.zeros((P, 3)), magnetisation=4e5, damping=0.1, temperature=300.0,
                field_shape=field_shape, field_amplitude=0.02 / MU0, field_frequency=3e5,
                seeds=rng.integers(0, 2 ** 31 - 1, size=M))


def oracle_member(inp, m, time_step, end_time, S, implicit=False, noise=None, rng="philox",
                  eps=1e-9, renorm=False):
    b, e = int(inp["member_offsets"][m]), int(inp["member_offsets"][m + 1])
    fs = inp["field_shape"] if isinstance(inp["field_shape"], str) else inp["field_shape"][m]
    pick = lambda v: v if np.ndim(v) == 0 else v[m]  # noqa: E731
    nz = None
    if noise is not None:
        nz = np.ascontiguousarray(noise[:, b:e, :]).reshape(-1)
    return O.full_dynamics(inp["radius"][b:e], inp["anisotropy"][b:e],
                           inp["anisotropy_axis"][b:e], inp["magnetisation_direction"][b:e],
                           inp["location"][b:e], pick(inp["magnetisation"]),
                           pick(inp["damping"]), pick(inp["temperature"]), renorm, False,
                           implicit, eps, time_step, end_time, S, int(inp["seeds"][m]),
                           field_shape=fs, field_amplitude=pick(inp["field_amplitude"]),
                           field_frequency=pick(inp["field_frequency"]), rng=rng, noise=nz)


def _steps(dt, end, S, K=23e3):
    red = O.reduced_parameters([10e-9], [K], 4e5, 0.1, 300.0, dt, end)
    return int(O.sample_schedule(red["dt"], red["end"], S)[1])


def run_gpu(inp, time_step, end_time, S, **kw):
    args = dict(inp)
    return core.simulate_ensemble(time_step=time_step, end_time=end_time, max_samples=S,
                                  interactions=False, **args, **kw)


# ------------------------------------------------------------------ RNG
def test_wiener_stream_fp64_matches_oracle():
    _print_case("wiener_fp64")
    for seed, p, j0 in [(1, 0, 0), (2 ** 31 - 2, 7, 12345), (123456789012, 3, 2 ** 40 + 3)]:
        dev = core.wiener_normals(seed, p, j0, 4099, precision="fp64")
        ref = O.wiener_normals(seed, p, j0, 4099)
        # same Philox words; log/sqrt/sin/cos from two math libraries: <= a few ulp
        assert np.allclose(dev, ref, rtol=1e-14, atol=1e-14)


def test_wiener_stream_fp32_matches_oracle():
    _print_case("wiener_fp32")
    dev = core.wiener_normals(99, 5, 0, 100000, precision="fp32")
    ref = O.wiener_normals(99, 5, 0, 100000)
    # Native FP32 log2/sqrt/sin/cos on u = (x + 1/2) 2^-32 rounded to FP32.
    # Typical error is a few FP32 ulp; the worst cases are draws with u next
    # to 1 (small Box-Muller radius), where FP32 keeps few bits of 1 - u:
    # absolute error stays below 2e-5, far under one step's FP32 state rounding
    # after scaling by sr sqrt(dt) ~ 1e-3.
    err = np.abs(dev - ref)
    assert np.median(err) < 2e-7
    assert np.max(err) < 2e-5


# ------------------------------------------------------------------ Heun, FP64
@pytest.mark.parametrize("shape", ["sine", "square", "constant"])
def test_heun_fp64_array_noise_matches_oracle(shape):
    _print_case("heun_fp64_array", shape=shape)
    M, dt, S = 6, 1e-14, 41
    end = 2000 * dt
    inp = ensemble_inputs(M, seed=1, field_shape=shape)
    inp["field_frequency"] = 3e8  # several field periods within the run
    n_steps = _steps(dt, end, S)
    noise = np.random.default_rng(7).normal(size=(n_steps, M, 3))
    g = run_gpu(inp, dt, end, S, precision="fp64", noise=noise)
    for m in range(M):
        o = O.full_dynamics(inp["radius"][m:m + 1], inp["anisotropy"][m:m + 1],
                            inp["anisotropy_axis"][m:m + 1],
                            inp["magnetisation_direction"][m:m + 1], inp["location"][m:m + 1],
                            4e5, 0.1, 300.0, False, False, False, 1e-9, dt, end, S,
                            int(inp["seeds"][m]), field_shape=shape,
                            field_amplitude=inp["field_amplitude"], field_frequency=3e8,
                            rng="array", noise=np.ascontiguousarray(noise[:, m, :]).reshape(-1))
        assert g["n_steps"] == o["stats"].n_steps
        assert np.array_equal(g["time"], o["time"])
        assert np.allclose(g["field"], o["field"], rtol=1e-13, atol=1e-9)
        for c in "xyz":
            assert np.allclose(g[c][m], o[c][0], rtol=0, atol=4e5 * 1e-11), (m, c)


def test_heun_fp64_philox_matches_oracle():
    _print_case("heun_fp64_philox")
    M, dt, S = 8, 1e-14, 21
    end = 3000 * dt
    inp = ensemble_inputs(M, seed=2)
    g = run_gpu(inp, dt, end, S, precision="fp64")
    for m in range(M):
        o = oracle_member(inp, m, dt, end, S)
        for c in "xyz":
            assert np.allclose(g[c][m], o[c][0], rtol=0, atol=4e5 * 1e-11), (m, c)


def test_heun_clusters_without_interactions_match_oracle():
    # members of several particles, interactions off: per-particle streams keyed
    # by (member seed, particle index in member)
    _print_case("heun_clusters_nointeract")
    M, n_per, dt, S = 3, 4, 1e-14, 11
    end = 1500 * dt
    inp = ensemble_inputs(M, seed=3, n_per=n_per)
    inp["location"] = np.random.default_rng(1).normal(size=(M * n_per, 3)) * 1e-8
    g = run_gpu(inp, dt, end, S, precision="fp64")
    for m in range(M):
        o = oracle_member(inp, m, dt, end, S)
        for c in "xyz":
            assert np.allclose(g[c][m * n_per:(m + 1) * n_per], o[c], rtol=0,
                               atol=4e5 * 1e-11), (m, c)


def test_heun_fp64_renorm_matches_oracle():
    _print_case("heun_fp64_renorm")
    M, dt, S = 4, 1e-14, 11
    end = 1000 * dt
    inp = ensemble_inputs(M, seed=4)
    g = run_gpu(inp, dt, end, S, precision="fp64", renorm=True)
    for m in range(M):
        o = oracle_member(inp, m, dt, end, S, renorm=True)
        for c in "xyz":
            assert np.allclose(g[c][m], o[c][0], rtol=0, atol=4e5 * 1e-11)
    mag = np.sqrt(g["final_state"][:, 0] ** 2 + g["final_state"][:, 1] ** 2
                  + g["final_state"][:, 2] ** 2)
    assert np.allclose(mag, 1.0, atol=1e-14)


def test_members_with_different_schedules():
    # different K per member -> different reduced dt, possibly different sample steps
    _print_case("heun_schedule_groups")
    M, dt, S = 5, 1e-14, 17
    end = 777 * dt
    inp = ensemble_inputs(M, seed=5, K=np.array([23e3, 23e3, 11e3, 40e3, 11e3]))
    g = run_gpu(inp, dt, end, S, precision="fp64", member_times=True)
    for m in range(M):
        o = oracle_member(inp, m, dt, end, S)
        assert np.array_equal(g["member_time"][m], o["time"])
        for c in "xyz":
            assert np.allclose(g[c][m], o[c][0], rtol=0, atol=4e5 * 1e-11)


@pytest.mark.parametrize("implicit", [False, True])
def test_members_with_different_steps_and_one_schedule(implicit):
    # K and alpha moved by 1e-6 (a finite-difference ensemble): every member has
    # its own reduced step but the same integer sample schedule, so the FP64
    # run packs them into one launch; each equals its run alone, bit for bit
    _print_case("heun_merged_schedules", implicit=implicit)
    M, dt, S = 6, 1e-14, 17
    end = 777.3 * dt      # no sample on an exact step: the schedules cannot split
    h = np.array([0, 1, -1, 0, 0, 1]) * 1e-6
    inp = ensemble_inputs(M, seed=8, K=23e3 * (1 + h))
    inp["damping"] = 0.1 * (1 + np.array([0, 0, 0, 1, -1, -1]) * 1e-6)
    kw = dict(precision="fp64", use_implicit=implicit, member_times=True)
    g = run_gpu(inp, dt, end, S, **kw)
    for m in range(M):
        o = oracle_member(inp, m, dt, end, S, implicit=implicit)
        assert np.array_equal(g["member_time"][m], o["time"])
        for c in "xyz":
            assert np.allclose(g[c][m], o[c][0], rtol=0, atol=4e5 * 1e-11), (m, c)
        one = {k: (v[m:m + 1] if k in ("radius", "anisotropy", "anisotropy_axis",
                                       "magnetisation_direction", "location", "seeds")
                   else v) for k, v in inp.items()}
        one["member_offsets"] = np.array([0, 1], dtype=np.int64)
        one["damping"] = inp["damping"][m]
        g1 = run_gpu(one, dt, end, S, **kw)
        for c in "xyz":
            np.testing.assert_array_equal(g1[c][0], g[c][m])


def test_ensemble_sum_matches_trajectories():
    _print_case("ensemble_sum")
    M, dt, S = 300, 1e-14, 9
    end = 500 * dt
    inp = ensemble_inputs(M, seed=6)
    for prec in ("fp64", "fp32"):
        g = run_gpu(inp, dt, end, S, precision=prec)
        for d, c in enumerate("xyz"):
            assert np.allclose(g["ensemble_m"][:, d], g[c].sum(axis=0), rtol=1e-12, atol=1e-6)


# ------------------------------------------------------------------ FP32 vs FP64
def test_heun_fp32_tracks_fp64_on_the_same_stream():
    _print_case("heun_fp32_vs_fp64")
    M, dt, S = 256, 1e-14, 11
    end = 5000 * dt
    inp = ensemble_inputs(M, seed=8)
    g64 = run_gpu(inp, dt, end, S, precision="fp64")
    g32 = run_gpu(inp, dt, end, S, precision="fp32")
    # FP32 rounding (~6e-8 per step) accumulating over 5000 steps
    for c in "xyz":
        assert np.max(np.abs(g32[c] - g64[c])) / 4e5 < 5e-5
    assert np.all(g32["status"] == 0) and np.all(g64["status"] == 0)


def test_fp32_run_is_deterministic():
    _print_case("fp32_deterministic")
    inp = ensemble_inputs(4096, seed=9)
    a = run_gpu(inp, 1e-14, 2000e-14, 5, precision="fp32")
    b = run_gpu(inp, 1e-14, 2000e-14, 5, precision="fp32")
    assert np.array_equal(a["final_state"], b["final_state"])
    assert np.array_equal(a["ensemble_m"], b["ensemble_m"])


# ------------------------------------------------------------------ implicit midpoint
def test_implicit_fp64_array_noise_matches_oracle():
    _print_case("implicit_fp64_array")
    M, dt, S = 4, 5e-13, 11
    end = 100 * dt
    inp = ensemble_inputs(M, seed=10)
    n_steps = _steps(dt, end, S)
    noise = np.random.default_rng(11).normal(size=(n_steps, M, 3))
    g = run_gpu(inp, dt, end, S, use_implicit=True, noise=noise)
    iters = 0
    for m in range(M):
        o = oracle_member(inp, m, dt, end, S, implicit=True, noise=noise, rng="array")
        iters += o["stats"].newton_iters
        # converged Newton roots at relative tolerance 1e-9
        for c in "xyz":
            assert np.allclose(g[c][m], o[c][0], rtol=0, atol=4e5 * 1e-8), (m, c)
    assert g["newton_failures"] == 0
    assert abs(g["newton_iters"] - iters) <= 0.02 * iters


def test_implicit_philox_matches_oracle():
    _print_case("implicit_fp64_philox")
    M, dt, S = 3, 5e-13, 6
    end = 60 * dt
    inp = ensemble_inputs(M, seed=12)
    g = run_gpu(inp, dt, end, S, use_implicit=True)
    for m in range(M):
        o = oracle_member(inp, m, dt, end, S, implicit=True)
        for c in "xyz":
            assert np.allclose(g[c][m], o[c][0], rtol=0, atol=4e5 * 1e-8)


# ------------------------------------------------------------------ API level
def test_model_simulate_heun_matches_oracle():
    _print_case("model_simulate")
    from magpy_amd import Model
    m = Model(radius=[12e-9], anisotropy=[23e3], anisotropy_axis=[[0., 0., 1.]],
              magnetisation_direction=[[0., 0., 1.]], location=[[0., 0., 0.]],
              magnetisation=400e3, damping=0.1, temperature=300.)
    res = m.simulate(end_time=1e-11, time_step=1e-14, max_samples=11, seed=1001,
                     implicit_solve=False)
    o = O.full_dynamics([12e-9], [23e3], [[0, 0, 1]], [[0, 0, 1]], [[0, 0, 0]], 400e3, 0.1, 300.,
                        False, True, False, 1e-9, 1e-14, 1e-11, 11, 1001, rng="philox")
    assert res.N == 1 and np.array_equal(res.time, o["time"])
    assert np.allclose(res.z[0], o["z"][0], rtol=0, atol=4e5 * 1e-11)


def test_ensemble_model_seeds_and_mean():
    _print_case("ensemble_model")
    from magpy_amd import EnsembleModel, Model
    base = Model(radius=[10e-9], anisotropy=[23e3], anisotropy_axis=[[0., 0., 1.]],
                 magnetisation_direction=[[0., 0., 1.]], location=[[0., 0., 0.]],
                 magnetisation=400e3, damping=0.1, temperature=300., field_shape='sine',
                 field_frequency=3e8, field_amplitude=0.02 / MU0)
    ens = EnsembleModel(16, base, radius=[[10e-9 * (1 + 0.01 * i)] for i in range(16)])
    er = ens.simulate(end_time=2e-12, time_step=1e-14, max_samples=21, random_state=3,
                      implicit_solve=False)
    np.random.seed(3)
    seeds = np.random.randint(np.iinfo(np.int32).max, size=16)
    for i in (0, 7, 15):
        o = O.full_dynamics([10e-9 * (1 + 0.01 * i)], [23e3], [[0, 0, 1]], [[0, 0, 1]],
                            [[0, 0, 0]], 400e3, 0.1, 300., False, True, False, 1e-9, 1e-14,
                            2e-12, 21, int(seeds[i]), field_shape='sine',
                            field_amplitude=0.02 / MU0, field_frequency=3e8, rng="philox")
        assert np.allclose(er.results[i].z[0], o["z"][0], rtol=0, atol=4e5 * 1e-11)
    mean = np.mean([r.magnetisation('z') for r in er.results], axis=0)
    assert np.allclose(er.ensemble_magnetisation('z'), mean, rtol=1e-12)
    assert np.isfinite(er.final_cycle_energy_dissipated(3e8))
