"""Implicit midpoint for interacting clusters (the reference's default
Model.simulate path, magpy/model.py:70) on the device against the oracle.

The oracle runs the reference's 3N x 3N Newton system per cluster step
(lib/integrators.cpp:576-651 with lib/optimisation.cpp:81-149 and the
multi-particle Jacobians of lib/llg.cpp:378-481); the device solves its
3 x 3 diagonal blocks with cluster-wide norms and dipolar F.  FP64 samples
agree to 1e-10 of Ms, Newton iteration counts are equal member by member,
for the Philox stream and for caller-given increments (the reference's
RngArray).  Caller-given increments also drive the Heun cluster kernel.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

core = pytest.importorskip("magpy_amd.core")
from magpy_amd import geometry as G  # noqa: E402

MU0 = 1.25663706e-6


def clusters(sizes, seed, phi=0.10):
    rng = np.random.default_rng(seed)
    e = G.cluster_ensemble(sizes, 10e-9, phi, 2.2, rng)
    M = len(sizes)
    e.update(magnetisation=4e5, damping=0.1, temperature=300.0, field_shape="sine",
             field_amplitude=0.02 / MU0, field_frequency=3e8,
             seeds=rng.integers(0, 2 ** 31 - 1, size=M))
    return e


def member(e, m):
    b, x = int(e["member_offsets"][m]), int(e["member_offsets"][m + 1])
    return b, x


def oracle_member(e, m, dt, end, S, implicit, noise=None, interactions=True):
    b, x = member(e, m)
    kw = dict(rng="philox")
    if noise is not None:
        kw = dict(rng="array", noise=np.ascontiguousarray(noise[:, b:x, :]).reshape(-1))
    return O.full_dynamics(e["radius"][b:x], e["anisotropy"][b:x], e["anisotropy_axis"][b:x],
                           e["magnetisation_direction"][b:x], e["location"][b:x], 4e5, 0.1,
                           300.0, False, interactions, implicit, 1e-9, dt, end, S,
                           int(e["seeds"][m]), field_shape="sine", field_amplitude=0.02 / MU0,
                           field_frequency=3e8, **kw)


def single(e, m):
    b, x = member(e, m)
    one = {k: (v[b:x] if k in ("radius", "anisotropy", "anisotropy_axis",
                               "magnetisation_direction", "location") else v)
           for k, v in e.items()}
    one["member_offsets"] = np.array([0, x - b])
    one["seeds"] = e["seeds"][m:m + 1]
    return one


@pytest.mark.parametrize("noise_kind", ["philox", "array"])
def test_implicit_clusters_match_oracle(noise_kind):
    print("CASE implicit_clusters", noise_kind, flush=True)
    sizes = [2, 3, 4, 5, 6, 7, 8, 1]
    e = clusters(sizes, 21)
    dt, S, n_steps = 5e-13, 9, 160
    end = n_steps * dt
    noise = None
    if noise_kind == "array":
        P = int(e["member_offsets"][-1])
        noise = np.random.default_rng(5).normal(size=(n_steps + 1, P, 3))
    g = core.simulate_ensemble(time_step=dt, end_time=end, max_samples=S, interactions=True,
                               use_implicit=True, precision="fp64", noise=noise, **e)
    assert np.all(g["status"] == 0) and g["newton_failures"] == 0
    total = 0
    for m in range(len(sizes)):
        o = oracle_member(e, m, dt, end, S, True, noise)
        b, x = member(e, m)
        for c in "xyz":
            assert np.allclose(g[c][b:x], o[c], rtol=0, atol=4e5 * 1e-10), (m, c)
        total += o["stats"].newton_iters
        assert o["stats"].newton_failures == 0
    assert g["newton_iters"] == total, (g["newton_iters"], total)
    # member by member: the device's count for one cluster alone
    for m in (1, 6):
        one = single(e, m)
        nz = None if noise is None else np.ascontiguousarray(noise[:, member(e, m)[0]:member(e, m)[1]])
        g1 = core.simulate_ensemble(time_step=dt, end_time=end, max_samples=S, interactions=True,
                                    use_implicit=True, precision="fp64", noise=nz, **one)
        o = oracle_member(e, m, dt, end, S, True, noise)
        assert g1["newton_iters"] == o["stats"].newton_iters
        b, x = member(e, m)
        for c in "xyz":
            np.testing.assert_array_equal(g1[c], g[c][b:x])


def test_implicit_interactions_change_the_trajectory():
    print("CASE implicit_clusters_interactions", flush=True)
    e = clusters([6, 3], 22, phi=0.3)
    dt, S = 5e-13, 5
    end = 200 * dt
    on = core.simulate_ensemble(time_step=dt, end_time=end, max_samples=S, interactions=True,
                                use_implicit=True, precision="fp64", **e)
    off = core.simulate_ensemble(time_step=dt, end_time=end, max_samples=S, interactions=False,
                                 use_implicit=True, precision="fp64", **e)
    assert np.max(np.abs(on["z"] - off["z"])) > 1e-6 * 4e5
    o_off = oracle_member(e, 0, dt, end, S, True, interactions=False)
    assert np.allclose(off["z"][:6], o_off["z"], rtol=0, atol=4e5 * 1e-10)


def test_implicit_clusters_packed_among_many():
    # uneven clusters up to 64 particles in one launch: each equals its run alone
    print("CASE implicit_clusters_packed", flush=True)
    rng = np.random.default_rng(23)
    sizes = list(rng.integers(1, 65, size=60))
    e = clusters(sizes, 24)
    dt, S = 5e-13, 4
    end = 30 * dt
    g = core.simulate_ensemble(time_step=dt, end_time=end, max_samples=S, interactions=True,
                               use_implicit=True, precision="fp64", **e)
    assert np.all(g["status"] == 0)
    for m in (0, 31, 59):
        g1 = core.simulate_ensemble(time_step=dt, end_time=end, max_samples=S,
                                    interactions=True, use_implicit=True, precision="fp64",
                                    **single(e, m))
        b, x = member(e, m)
        for c in "xyz":
            np.testing.assert_array_equal(g1[c], g[c][b:x])
    # one 17-particle cluster against the oracle
    m = int(np.argmin(np.abs(np.array(sizes) - 17)))
    o = oracle_member(e, m, dt, end, S, True)
    b, x = member(e, m)
    for c in "xyz":
        assert np.allclose(g[c][b:x], o[c], rtol=0, atol=4e5 * 1e-10)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_heun_clusters_array_noise_match_oracle(precision):
    print("CASE heun_clusters_array_noise", precision, flush=True)
    sizes = [2, 5, 9, 16, 1]
    e = clusters(sizes, 25)
    dt, S, n_steps = 1e-14, 6, 700
    end = n_steps * dt
    P = int(e["member_offsets"][-1])
    noise = np.random.default_rng(6).normal(size=(n_steps + 1, P, 3))
    g = core.simulate_ensemble(time_step=dt, end_time=end, max_samples=S, interactions=True,
                               precision=precision, noise=noise, **e)
    tol = 1e-10 if precision == "fp64" else 2e-5
    for m in range(len(sizes)):
        o = oracle_member(e, m, dt, end, S, False, noise)
        b, x = member(e, m)
        for c in "xyz":
            assert np.allclose(g[c][b:x], o[c], rtol=0, atol=4e5 * tol), (m, c)
    # the increments are the ones given: another array moves the result
    g2 = core.simulate_ensemble(time_step=dt, end_time=end, max_samples=S, interactions=True,
                                precision=precision, noise=-noise, **e)
    assert np.max(np.abs(g2["z"] - g["z"])) > 1e-5 * 4e5


def test_model_simulate_with_the_reference_defaults():
    # Model.simulate's defaults are interactions=True, implicit_solve=True
    # (magpy/model.py:70): a dimer through the Python API
    print("CASE implicit_model_defaults", flush=True)
    from magpy_amd import Model
    axes = [[0.0, 0.6, 0.8], [0.8, 0.0, 0.6]]
    loc = [[0.0, 0.0, 0.0], [0.0, 0.0, 24e-9]]
    m = Model(radius=[10e-9, 11e-9], anisotropy=[23e3, 20e3], anisotropy_axis=axes,
              magnetisation_direction=axes, location=loc, magnetisation=4e5, damping=0.1,
              temperature=300.0, field_shape='sine', field_frequency=3e8,
              field_amplitude=0.02 / MU0)
    res = m.simulate(end_time=100 * 5e-13, time_step=5e-13, max_samples=11, seed=99)
    o = O.full_dynamics([10e-9, 11e-9], [23e3, 20e3], axes, axes, loc, 4e5, 0.1, 300.0, False,
                        True, True, 1e-9, 5e-13, 100 * 5e-13, 11, 99, field_shape="sine",
                        field_amplitude=0.02 / MU0, field_frequency=3e8, rng="philox")
    assert res.N == 2
    for i in range(2):
        assert np.allclose(res.z[i], o["z"][i], rtol=0, atol=4e5 * 1e-10)
        assert np.allclose(res.x[i], o["x"][i], rtol=0, atol=4e5 * 1e-10)
