"""Sensitivities, host side: what pins each piece without a GPU.

1. The product's reduced parameters (dt, sr sqrt(dt), alpha, happ) equal the
   oracle's SI -> reduced conversion (lib/simulation.cpp:501-549).
2. The product's analytic dq/dtheta equals central differences of the
   oracle's conversion.
3. The oracle's forward-mode tangent of the Heun path equals central
   differences of the oracle's own run (FP64; a dropped term or a wrong
   sign in the tangent fails this).
4. The oracle's path (fused Heun form, SI field phase) equals the
   reference-structured Heun of orc_full_dynamics (lib/integrators.cpp:
   372-405 with dense diffusion) on the same Philox increments.
5. Argument checks of mp_adjoint_create run before any device call.
"""
import numpy as np
import pytest

from oracle import oracle as O

core = pytest.importorskip("magpy_amd.core")

MS, T, DT, F = 4e5, 300.0, 1e-14, 3e5
THETA = np.array([23e3, 0.1, 0.02 / 1.25663706e-6, 10e-9])


def oracle_q(theta, shape):
    """q [P, 4] from the oracle's reduction, one particle per member."""
    K, al, H, rm = theta
    q = np.zeros((len(shape), 4))
    for p, s in enumerate(shape):
        red = O.reduced_parameters([rm * s], [K], MS, al, T, DT, 1e-9, H, F)
        q[p] = [red["dt"], red["sr"][0] * np.sqrt(red["dt"]), al, red["happ"]]
    return q


def test_reduced_parameters_match_oracle():
    shape = np.exp(0.1 * np.random.default_rng(0).normal(size=7))
    q, dq = core.adjoint_reduced(THETA, shape, MS, T, DT)
    qo = oracle_q(THETA, shape)
    assert np.allclose(q, qo, rtol=1e-14, atol=0)


@pytest.mark.parametrize("i", range(4))
def test_parameter_derivatives_match_differences(i):
    shape = np.exp(0.1 * np.random.default_rng(1).normal(size=5))
    _, dq = core.adjoint_reduced(THETA, shape, MS, T, DT)
    h = 1e-5 * THETA[i]
    tp, tm = THETA.copy(), THETA.copy()
    tp[i] += h
    tm[i] -= h
    fd = (oracle_q(tp, shape) - oracle_q(tm, shape)) / (2 * h)
    # natural scale of dq_j / dtheta_i: |q_j| / |theta_i| (1/H_k for happ vs H)
    q = oracle_q(THETA, shape)
    scale = np.max(np.abs(q), axis=0, keepdims=True) / abs(THETA[i])
    err = np.abs(dq[:, :, i] - fd) / scale
    assert np.all(err < 1e-7), (i, err)
    # sig = sr sqrt(dt) does not depend on K
    if i == 0:
        assert np.all(dq[:, 1, 0] == 0.0)


def small_problem(P=4, n_steps=300, seed=3):
    rng = np.random.default_rng(seed)
    ax = rng.normal(size=(P, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    red = O.reduced_parameters([10e-9], [23e3], MS, 0.1, T, DT, 1e-9, 15915.5, F)
    q = dict(dt=red["dt"], sig=red["sr"][0] * np.sqrt(red["dt"]) * np.exp(0.2 * rng.normal(size=P)),
             alpha=0.1, happ=red["happ"])
    ks = [0, 7, 100, 299, n_steps]
    return ax, q, ks, rng.integers(1, 2 ** 40, size=P)


def test_oracle_tangent_matches_differences():
    P, N = 4, 300
    ax, q, ks, seeds = small_problem(P, N)
    fdt = F * DT * 1e4          # a fast field so the phase term matters

    def run(qq):
        pr = O.path_problem(ax, ax, qq["dt"], qq["sig"], qq["alpha"], qq["happ"], fdt, seeds, N, ks)
        return O.path_run(pr)

    dq = np.zeros((P, 4, 4))
    for j in range(4):
        dq[:, j, j] = 1.0
    pr = O.path_problem(ax, ax, q["dt"], q["sig"], q["alpha"], q["happ"], fdt, seeds, N, ks)
    tan, tan_f = O.path_tangent(pr, dq)
    for j, name in enumerate(("dt", "sig", "alpha", "happ")):
        h = 1e-4 * np.max(np.abs(q[name]))
        qp, qm = dict(q), dict(q)
        qp[name] = q[name] + h
        qm[name] = q[name] - h
        (a, af, _), (b, bf, _) = run(qp), run(qm)
        fd, fdf = (a - b) / (2 * h), (af - bf) / (2 * h)
        scale = np.max(np.abs(tan[:, j])) + 1e-300
        assert np.max(np.abs(fd - tan[:, j])) < 1e-6 * scale, name
        assert abs(fdf - tan_f[j]) < 1e-6 * scale, name
    # the sample at step 0 has no sensitivity; the final sample equals the final state
    assert np.all(tan[0] == 0.0)
    assert np.allclose(tan[-1], tan_f, rtol=0, atol=0)


def test_oracle_path_is_the_reference_heun():
    """Path run == orc_full_dynamics (reduced Heun with dense diffusion,
    field sin(2 pi f_red t_red)) for one-particle members on the Philox
    stream, at the reference's own sample steps."""
    rng = np.random.default_rng(4)
    P, S = 3, 9
    ax = rng.normal(size=(P, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    radius = 10e-9 * np.exp(0.1 * rng.normal(size=P))
    seeds = rng.integers(1, 2 ** 31 - 1, size=P)
    H, end = 0.02 / 1.25663706e-6, 2000 * DT
    mz_ref = []
    for p in range(P):
        o = O.full_dynamics([radius[p]], [23e3], ax[p], ax[p], [[0, 0, 0]], MS, 0.1, T, False,
                            False, False, 1e-9, DT, end, S, int(seeds[p]), field_shape="sine",
                            field_amplitude=H, field_frequency=F, rng="philox")
        mz_ref.append(o["z"][0] / MS)
    red = [O.reduced_parameters([r], [23e3], MS, 0.1, T, DT, end, H, F) for r in radius]
    ks, n = O.sample_schedule(red[0]["dt"], red[0]["end"], S)
    sig = [r["sr"][0] * np.sqrt(r["dt"]) for r in red]
    pr = O.path_problem(ax, ax, red[0]["dt"], sig, 0.1, red[0]["happ"], F * DT, seeds, n, ks)
    mz, _, _ = O.path_run(pr)
    assert np.allclose(mz, np.sum(mz_ref, axis=0), rtol=0, atol=1e-11)


def test_create_rejects_bad_requests():
    P = 2
    base = dict(radius_shape=np.ones(P), anisotropy_axis=np.tile([0, 0, 1.0], (P, 1)),
                magnetisation_direction=np.tile([0, 0, 1.0], (P, 1)), seeds=np.arange(P),
                magnetisation=MS, temperature=T, field_frequency=F, time_step=DT, n_steps=100,
                sample_steps=[0, 50, 100])
    for bad in (dict(checkpoint_interval=10), dict(checkpoint_interval=0),
                dict(sample_steps=[0, 101]), dict(sample_steps=[50, 10]), dict(n_steps=0),
                dict(radius_shape=np.array([1.0, -1.0]))):
        kw = dict(base)
        kw.update(bad)
        with pytest.raises(ValueError):
            core.Adjoint(**kw)
    with pytest.raises(ValueError):
        core.adjoint_reduced([23e3, 0.0, 1.0, 1e-8], np.ones(2), MS, T, DT)
