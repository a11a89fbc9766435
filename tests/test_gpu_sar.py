"""SAR (final-cycle hysteresis energy) from device runs against the oracle.

EnsembleModel.simulate reduces the ensemble magnetisation on the device in
FP64 (trajectories=False keeps nothing else); EnsembleResults then applies
the reference's estimator (magpy/results.py:167-217), whose closed form is
pinned on the CPU (tests/test_results_cpu.py).  The oracle runs every member
with the same Philox increments and its Results go through the same
estimator.  FP64: the energies agree to 1e-10 of mu0 H0 Ms (the loop's
natural scale; trajectories differ only by operation order).  FP32 state:
within 1e-4 of mu0 H0 Ms over two cycles (66,667 steps; 4.8e-6 measured); the FP32 error of
the full AC-field workload against FP64 is reported with the benchmark.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

core = pytest.importorskip("magpy_amd.core")
from magpy_amd import EnsembleModel, Model  # noqa: E402
from magpy_amd.results import EnsembleResults, Results  # noqa: E402

MU0 = 1.25663706e-6
MS, H0, FREQ, DT = 4e5, 0.02 / 1.25663706e-6, 3e9, 1e-14
M, CYCLES, PER_CYCLE, SEED = 16, 2, 400, 5


def ensemble():
    rng = np.random.default_rng(17)
    r = 10e-9 * np.exp(0.1 * rng.normal(size=M))
    axes = rng.normal(size=(M, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    base = Model(radius=[10e-9], anisotropy=[23e3], anisotropy_axis=[[0., 0., 1.]],
                 magnetisation_direction=[[0., 0., 1.]], location=[[0., 0., 0.]],
                 magnetisation=MS, damping=0.1, temperature=300., field_shape='sine',
                 field_frequency=FREQ, field_amplitude=H0)
    ens = EnsembleModel(M, base, radius=[[x] for x in r], anisotropy_axis=[[a] for a in axes],
                        magnetisation_direction=[[a] for a in axes])
    return ens, r, axes


def oracle_results(r, axes, end, S):
    np.random.seed(SEED)   # the reference's seed derivation (magpy/model.py:202-203)
    seeds = np.random.randint(np.iinfo(np.int32).max, size=M)
    out = []
    for i in range(M):
        o = O.full_dynamics([r[i]], [23e3], [axes[i]], [axes[i]], [[0, 0, 0]], MS, 0.1, 300.,
                            False, True, False, 1e-9, DT, end, S, int(seeds[i]),
                            field_shape='sine', field_amplitude=H0, field_frequency=FREQ,
                            rng='philox')
        out.append(Results(o['time'], o['field'], {0: o['x'][0]}, {0: o['y'][0]},
                           {0: o['z'][0]}, 1))
    return EnsembleResults(out)


def test_device_sar_matches_oracle_estimator():
    print("CASE sar_fp64_fp32_vs_oracle", flush=True)
    ens, r, axes = ensemble()
    end = CYCLES / FREQ
    S = CYCLES * PER_CYCLE + 1
    kw = dict(end_time=end, time_step=DT, max_samples=S, random_state=SEED,
              implicit_solve=False)
    dev = ens.simulate(precision='fp64', trajectories=False, **kw)
    orc = oracle_results(r, axes, end, S)
    np.testing.assert_array_equal(dev.time, orc.time)
    np.testing.assert_array_equal(dev.field, orc.field)
    np.testing.assert_allclose(dev.ensemble_magnetisation('z'), orc.ensemble_magnetisation('z'),
                               rtol=0, atol=1e-10 * MS)
    scale = MU0 * H0 * MS
    e_dev = dev.final_cycle_energy_dissipated(FREQ)
    e_orc = orc.final_cycle_energy_dissipated(FREQ)
    print("final-cycle energy: device %.12e oracle %.12e (scale %.3e)" % (e_dev, e_orc, scale),
          flush=True)
    assert abs(e_orc) > 1e-3 * scale          # a real loop, not a rounding residue
    assert abs(e_dev - e_orc) < 1e-10 * scale
    assert abs(dev.energy_dissipated() - orc.energy_dissipated()) < 1e-10 * scale
    # kept trajectories give the same device sums
    full = ens.simulate(precision='fp64', trajectories=True, **kw)
    assert full.final_cycle_energy_dissipated(FREQ) == e_dev
    # FP32 state
    e32 = ens.simulate(precision='fp32', trajectories=False, **kw).final_cycle_energy_dissipated(
        FREQ)
    print("fp32 final-cycle energy %.8e (%.3e of scale from FP64)" % (e32, abs(e32 - e_dev) / scale),
          flush=True)
    assert abs(e32 - e_dev) < 1e-4 * scale
