"""Checkpoint and resume: a run split at any step and resumed from its state
equals the unsplit run bit for bit.

The Wiener increments of step n are a pure function of (seed, particle, n)
and the applied field of t_{n+1}, so nothing but the state carries over a
split.  Each path is checked: independent particles in FP32 (split inside a
4-step group and across a launch block) and FP64, the implicit midpoint
scheme, interacting clusters and thermal replicas on the product path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

core = pytest.importorskip("magpy_amd.core")
from magpy_amd import geometry as G  # noqa: E402

MU0 = 1.25663706e-6


def independent(P, seed):
    rng = np.random.default_rng(seed)
    axes = G.uniform_random_axes(P, rng)
    return dict(member_offsets=np.arange(P + 1, dtype=np.int64),
                radius=10e-9 * np.exp(0.1 * rng.normal(size=P)), anisotropy=np.full(P, 23e3),
                anisotropy_axis=axes, magnetisation_direction=axes.copy(),
                location=np.zeros((P, 3)), magnetisation=4e5, damping=0.1, temperature=300.0,
                field_shape="sine", field_amplitude=0.02 / MU0, field_frequency=3e8,
                seeds=rng.integers(0, 2 ** 31 - 1, size=P))


def split_and_compare(e, dt, n_steps, split, S, **kw):
    end = n_steps * dt
    full = core.simulate_ensemble(time_step=dt, end_time=end, max_samples=S, **e, **kw)
    # first leg: a session advanced to the split step
    ses = core.Session(time_step=dt, end_time=end, max_samples=S, **e, **kw)
    done = ses.advance(split)
    assert done == split
    state, step = ses.state()
    assert step == split
    second = core.simulate_ensemble(time_step=dt, end_time=end, max_samples=S, start_step=split,
                                    start_state=state, **e, **kw)
    assert np.array_equal(second["final_state"], full["final_state"])
    assert np.array_equal(second["status"], full["status"])
    # samples taken before the split are NaN (a prefix of the schedule, not
    # empty since split > 0); every later sample agrees
    later = ~np.isnan(second["x"][0])
    first = int(np.argmax(later))
    assert 0 < first and later[first:].all()
    for c in "xyz":
        assert np.array_equal(second[c][:, later], full[c][:, later]), c
        assert np.all(np.isnan(second[c][:, ~later])), c
    assert np.array_equal(second["ensemble_m"][later], full["ensemble_m"][later])
    return full


@pytest.mark.parametrize("precision,split", [("fp32", 4097), ("fp32", 2048), ("fp64", 1001)])
def test_independent_resume(precision, split, monkeypatch):
    print("CASE resume_independent", precision, split, flush=True)
    # short launch blocks so the split falls inside and across blocks
    monkeypatch.setenv("MAGPY_AMD_STEPS_PER_LAUNCH", "1024")
    split_and_compare(independent(300, 1), 1e-14, 5000, split, 11, precision=precision)


def test_implicit_resume():
    print("CASE resume_implicit", flush=True)
    split_and_compare(independent(64, 2), 1e-13, 300, 131, 7, precision="fp64",
                      use_implicit=True)


def test_cluster_resume():
    print("CASE resume_clusters", flush=True)
    rng = np.random.default_rng(3)
    sizes = list(rng.integers(2, 40, size=50))
    e = G.cluster_ensemble(sizes, 10e-9, 0.1, 2.2, rng)
    e.update(magnetisation=4e5, damping=0.1, temperature=300.0, field_shape="sine",
             field_amplitude=0.02 / MU0, field_frequency=3e8,
             seeds=rng.integers(0, 2 ** 31 - 1, size=len(sizes)))
    split_and_compare(e, 1e-14, 1000, 377, 6, precision="fp32", interactions=True)


def test_replica_resume():
    print("CASE resume_replicas", flush=True)
    rng = np.random.default_rng(4)
    n, reps = 32, 64
    loc = G.random_cluster(n, 10e-9, 0.15, 2.2, rng)
    axes = G.uniform_random_axes(n, rng)
    e = dict(member_offsets=np.arange(reps + 1, dtype=np.int64) * n,
             radius=np.full(n * reps, 10e-9), anisotropy=np.full(n * reps, 23e3),
             anisotropy_axis=np.tile(axes, (reps, 1)),
             magnetisation_direction=np.tile(axes, (reps, 1)), location=np.tile(loc, (reps, 1)),
             magnetisation=4e5, damping=0.1, temperature=300.0, field_shape="sine",
             field_amplitude=0.02 / MU0, field_frequency=3e8,
             seeds=rng.integers(0, 2 ** 31 - 1, size=reps))
    ses = core.Session(time_step=1e-14, end_time=1e-12, max_samples=3, interactions=True,
                       precision="fp32", **e)
    assert ses.info()["dipolar"] == "gemm_bf16x3"
    split_and_compare(e, 1e-14, 300, 101, 5, precision="fp32", interactions=True)
