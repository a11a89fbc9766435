"""Discrete orientation model on the device vs the FP64 oracle restatement
of simulation::dom_ensemble_dynamics (pinned by the reference's rk45 and
transition-matrix known answers and the constant-rate closed form)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

core = pytest.importorskip("magpy_amd.core")
import magpy_amd  # noqa: E402

MU0 = 1.25663706e-6


@pytest.mark.parametrize("shape,n_comp", [("sine", 1), ("square", 1), ("square_f", 5),
                                          ("constant", 1)])
def test_dom_matches_oracle(shape, n_comp):
    print("CASE dom", shape, flush=True)
    rng = np.random.default_rng(1)
    P, S = 96, 41
    r = 8e-9 * np.exp(0.15 * rng.normal(size=P))
    V = 4 / 3 * np.pi * r ** 3
    K, T, ms, al = 23e3, 300.0, 4e5, 0.1
    H, f = 0.02 / MU0, 3e5
    ip = np.stack([rng.uniform(size=P), np.zeros(P)], 1)
    ip[:, 1] = 1 - ip[:, 0]
    dt, end = 1e-12, 2 / f
    g = core.simulate_dom_ensemble(V, K, T, ms, al, ip, dt, end, S, shape, H, f, n_comp)
    assert np.all(g["status"] == 0)
    hk = 2 * K / MU0 / ms
    for p in (0, 17, 95):
        o = O.dom_ensemble_dynamics(V[p], K, T, ms, al, shape, H / hk, f, ip[p], dt, end, S,
                                    n_components=n_comp)
        # the RK45 controller's accept/reject and step sizes depend on the last
        # bits of exp / sin (device libm vs glibc); trajectories then agree to
        # the accumulated tolerance of the controller (tol = time_step per step)
        assert np.allclose(g["mz"][p], o["mz"], rtol=0, atol=1e-7), p
        assert np.allclose(g["time"], o["time"], rtol=1e-15, atol=0)
    assert np.allclose(g["field"] / hk, O.dom_ensemble_dynamics(
        V[0], K, T, ms, al, shape, H / hk, f, ip[0], dt, end, S, n_components=n_comp)["field"],
        rtol=0, atol=1e-12)


def test_domodel_api():
    print("CASE domodel", flush=True)
    m = magpy_amd.DOModel(9e-9, 23e3, [1.0, 0.0], 4e5, 0.1, 300.0, field_shape="sine",
                          field_frequency=3e5, field_amplitude=0.02 / MU0)
    res = m.simulate(end_time=1 / 3e5, time_step=1e-12, max_samples=21)
    assert res.N == 1 and res.z[0][0] == 1.0 and np.all(res.x[0] == 0)
    assert np.all(np.abs(res.z[0]) <= 1 + 1e-9)


def test_unintegrable_particle_is_flagged_not_hung():
    # a step budget of 10 ends the particle early with status 1 and NaN samples
    print("CASE dom_budget", flush=True)
    g = core.simulate_dom_ensemble([4e-25, 4e-24], 23e3, 300.0, 4e5, 0.1, [[1.0, 0.0]], 1e-12,
                                   1e-6, 11, "sine", 1e4, 3e5, 1, max_steps=10)
    assert np.all(g["status"] == 1)
    assert np.all(np.isnan(g["mz"][:, -1]))
