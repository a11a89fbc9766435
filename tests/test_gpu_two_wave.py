"""The two-wave Heun kernels equal the one-wave kernel bit for bit.

For ensembles of at most one integrating wave per SIMD the library runs
heun_indep_pc_kernel (one wave producing the noise into LDS, one
integrating): with the producer forming the normals (mode 1), or, in FP32,
writing the Philox words for the integrating wave to form them (mode 2);
larger ensembles run heun_indep_kernel (mode 0).  All compute every value
with the same expressions, so forcing each kernel (MAGPY_AMD_HEUN_TWO_WAVE)
on the same ensemble must give identical samples, final states and ensemble
sums: with a shared field table (every member on one field, FP32), with
per-lane fields (members on different frequencies and shapes), with samples
inside 4-step groups, launch blocks that do not end on a group, a particle
count that is not a multiple of 64, renormalisation, in FP32 and FP64.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

core = pytest.importorskip("magpy_amd.core")
from magpy_amd import geometry as G  # noqa: E402

MU0 = 1.25663706e-6


def ensemble(P, seed, mixed):
    rng = np.random.default_rng(seed)
    axes = G.uniform_random_axes(P, rng)
    e = dict(member_offsets=np.arange(P + 1, dtype=np.int64),
             radius=10e-9 * np.exp(0.1 * rng.normal(size=P)), anisotropy=np.full(P, 23e3),
             anisotropy_axis=axes, magnetisation_direction=axes.copy(),
             location=np.zeros((P, 3)), magnetisation=4e5, damping=0.1, temperature=300.0,
             field_shape="sine", field_amplitude=0.02 / MU0, field_frequency=3e8,
             seeds=rng.integers(0, 2 ** 31 - 1, size=P))
    if mixed:
        e["field_frequency"] = rng.choice([1e8, 3e8, 7e8], size=P)
        e["field_shape"] = rng.choice(["sine", "square", "constant"], size=P).tolist()
    return e


def run_modes(monkeypatch, precision, **kw):
    out = {}
    for tw in ("0", "1", "2"):
        monkeypatch.setenv("MAGPY_AMD_HEUN_TWO_WAVE", tw)
        out[tw] = core.simulate_ensemble(precision=precision, **kw)
    return out


@pytest.mark.parametrize("precision,P,mixed,renorm,block", [
    ("fp32", 4096, False, False, 0), ("fp32", 1000, True, False, 0),
    ("fp32", 777, False, True, 1001), ("fp32", 2048, True, True, 998),
    ("fp64", 4096, False, False, 0), ("fp64", 777, True, True, 1001)])
def test_two_wave_equals_one_wave(monkeypatch, precision, P, mixed, renorm, block):
    print("CASE two_wave", precision, P, mixed, renorm, block, flush=True)
    if block:
        monkeypatch.setenv("MAGPY_AMD_STEPS_PER_LAUNCH", str(block))
    e = ensemble(P, 17 + P, mixed)
    dt = 1e-14
    # 4,003 steps; 37 samples fall at steps that are not multiples of 4
    out = run_modes(monkeypatch, precision, time_step=dt, end_time=4002.5 * dt,
                    max_samples=37, renorm=renorm, **e)
    one = out["0"]
    assert np.all(one["status"] == 0)
    for mode in ("1", "2"):
        two = out[mode]
        assert one["n_steps"] == two["n_steps"]
        assert np.all(two["status"] == 0)
        for c in "xyz":
            assert np.array_equal(one[c], two[c]), (mode, c)
        assert np.array_equal(one["final_state"], two["final_state"]), mode
        assert np.array_equal(one["ensemble_m"], two["ensemble_m"]), mode
