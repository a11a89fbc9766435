"""Pathwise sensitivities on the device vs the FP64 oracle.

- FP64 forward pass == the oracle's path run (same Philox increments, same
  field phase): ensemble <M_z> at the samples and the final <m_z> to 1e-12.
- FP64 reverse pass == the oracle's forward-mode tangent (an independent
  derivation: product-rule directional derivatives, not transposes), chained
  to theta with the product's dq/dtheta (pinned on the CPU against
  differences of the oracle's reduction): relative 1e-9.
- Splitting the passes into many launches changes nothing but the order of
  the FP64 gradient sums.
- FP32 state and adjoint vs FP64: relative 1e-3 (the FP32 trajectory itself
  differs from FP64 by ~1e-6 relative per step, which the gradient inherits).
- torch.autograd through SARObjective vs central differences with common
  random numbers (FP64 forward passes, steps 1e-4 and 5e-5 relative,
  Richardson-extrapolated): relative 1e-6 for the FP64 adjoint and 1e-3 for
  the FP32 adjoint.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

core = pytest.importorskip("magpy_amd.core")
torch = pytest.importorskip("torch")
from magpy_amd import adjoint as A  # noqa: E402
from magpy_amd import geometry as G  # noqa: E402

MS, T, DT = 4e5, 300.0, 1e-14
THETA = np.array([23e3, 0.1, 0.02 / 1.25663706e-6, 10e-9])


def problem(P, seed):
    rng = np.random.default_rng(seed)
    ax = G.uniform_random_axes(P, rng)
    shape = np.exp(0.1 * rng.normal(size=P))
    seeds = rng.integers(1, 2 ** 31 - 1, size=P)
    return shape, ax, seeds


def runner(shape, ax, seeds, f, N, ks, C, precision):
    return core.Adjoint(shape, ax, ax, seeds, MS, T, f, DT, N, ks, C, precision)


def oracle_tangent(shape, ax, seeds, f, N, ks, theta):
    q, dq = core.adjoint_reduced(theta, shape, MS, T, DT)
    pr = O.path_problem(ax, ax, q[:, 0], q[:, 1], q[:, 2], q[:, 3], f * DT, seeds, N, ks)
    mz, mz_f, _ = O.path_run(pr)
    # directions = theta_i: dq[p, j, i] -> [P, D=4 (i), 4 (j)]
    tan, tan_f = O.path_tangent(pr, np.ascontiguousarray(np.transpose(dq, (0, 2, 1))))
    return mz, mz_f, tan, tan_f


CASE = dict(P=200, f=1e10, N=3001, ks=[0, 5, 1000, 2999, 3000, 3001], C=100)


def test_fp64_forward_matches_oracle():
    print("CASE adjoint_forward_fp64", flush=True)
    shape, ax, seeds = problem(CASE["P"], 1)
    r = runner(shape, ax, seeds, CASE["f"], CASE["N"], CASE["ks"], CASE["C"], "fp64")
    ens, fin, st = r.forward(THETA)
    assert st == 0
    mz, mz_f, _, _ = oracle_tangent(shape, ax, seeds, CASE["f"], CASE["N"], CASE["ks"], THETA)
    P = CASE["P"]
    assert np.allclose(ens, MS * mz / P, rtol=0, atol=MS * 1e-12)
    assert abs(fin - mz_f / P) < 1e-12


def test_fp64_gradient_matches_oracle_tangent():
    print("CASE adjoint_reverse_fp64", flush=True)
    shape, ax, seeds = problem(CASE["P"], 2)
    r = runner(shape, ax, seeds, CASE["f"], CASE["N"], CASE["ks"], CASE["C"], "fp64")
    r.forward(THETA)
    w = np.random.default_rng(3).normal(size=len(CASE["ks"]))
    wf = 0.7
    grad = r.reverse(w, wf)
    _, _, tan, tan_f = oracle_tangent(shape, ax, seeds, CASE["f"], CASE["N"], CASE["ks"], THETA)
    P = CASE["P"]
    ref = (MS / P) * (w @ tan) + (wf / P) * tan_f
    print("grad", grad, "oracle", ref, flush=True)
    assert np.all(np.abs(grad - ref) <= 1e-9 * np.abs(ref)), (grad, ref)


def test_split_launches_change_nothing(monkeypatch):
    print("CASE adjoint_split_launches", flush=True)
    shape, ax, seeds = problem(70, 4)     # a partial last wave
    ks, N = [0, 400, 401, 1203, 1203, 2002], 2002
    w = np.linspace(-1, 1, len(ks))
    r = runner(shape, ax, seeds, 1e10, N, ks, 100, "fp64")
    ens1, fin1, _ = r.forward(THETA)
    g1 = r.reverse(w, 0.3)
    monkeypatch.setenv("MAGPY_AMD_ADJOINT_STEPS_PER_LAUNCH", "200")
    r2 = runner(shape, ax, seeds, 1e10, N, ks, 100, "fp64")
    r2.begin(THETA)
    done = 0
    while done < N:
        done = r2.forward_advance(333)
    ens2, fin2, _ = r2.forward_result()
    r2.reverse_begin(w, 0.3)
    left = N
    while left > 0:
        left = r2.reverse_advance(450)
    g2 = r2.reverse_result()
    assert np.array_equal(ens1, ens2) and fin1 == fin2
    assert np.allclose(g1, g2, rtol=1e-12, atol=0)


def test_fp32_gradient_tracks_fp64():
    print("CASE adjoint_fp32_vs_fp64", flush=True)
    shape, ax, seeds = problem(CASE["P"], 5)
    w = np.random.default_rng(6).normal(size=len(CASE["ks"]))
    g = {}
    for prec in ("fp64", "fp32"):
        r = runner(shape, ax, seeds, CASE["f"], CASE["N"], CASE["ks"], CASE["C"], prec)
        r.forward(THETA)
        g[prec] = r.reverse(w, 0.5)
    err = np.abs(g["fp32"] - g["fp64"]) / np.abs(g["fp64"])
    print("fp32 vs fp64 relative", err, flush=True)
    assert np.all(err < 1e-3)


def sar_objective(precision, P=256, f=1e9, cycles=2, per_cycle=100):
    shape, ax, seeds = problem(P, 7)
    end = cycles / f
    return A.SARObjective(shape, ax, ax, seeds, MS, T, f, DT, end, cycles * per_cycle + 1,
                          checkpoint_interval=1000, precision=precision)


def test_autograd_matches_common_random_number_differences():
    print("CASE adjoint_autograd_vs_crn_fd", flush=True)
    obj64 = sar_objective("fp64")
    obj32 = sar_objective("fp32")
    grads = {}
    for name, obj in (("fp64", obj64), ("fp32", obj32)):
        for out in ("sar", "final_mz"):
            th = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
            res = obj(th)
            res[out].backward()
            grads[name, out] = th.grad.numpy().copy()

    def central(out, i, rel):
        h = rel * THETA[i]
        tp, tm = THETA.copy(), THETA.copy()
        tp[i] += h
        tm[i] -= h
        with torch.no_grad():
            a = float(obj64(torch.tensor(tp))[out])
            b = float(obj64(torch.tensor(tm))[out])
        return (a - b) / (2 * h)

    for out in ("sar", "final_mz"):
        # central differences at 1e-4 and 5e-5 relative, Richardson-extrapolated
        # (the O(h^2) term is ~1e-6 of the smallest components at h = 1e-4)
        fd = np.array([(4 * central(out, i, 5e-5) - central(out, i, 1e-4)) / 3
                       for i in range(4)])
        e64 = np.abs(grads["fp64", out] - fd) / np.abs(fd)
        e32 = np.abs(grads["fp32", out] - fd) / np.abs(fd)
        print(out, "fd", fd, "fp64 rel", e64, "fp32 rel", e32, flush=True)
        assert np.all(e64 < 1e-6), (out, e64)
        assert np.all(e32 < 1e-3), (out, e32)


def test_backward_after_another_forward_raises():
    print("CASE adjoint_stale_backward", flush=True)
    obj = sar_objective("fp32", P=64, f=1e10, cycles=1, per_cycle=20)
    th = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    res = obj(th)
    with torch.no_grad():
        obj(torch.tensor(THETA * 1.01))
    with pytest.raises(RuntimeError):
        res["sar"].backward()


def test_value_only_forward_has_no_reverse():
    print("CASE adjoint_value_only", flush=True)
    shape, ax, seeds = problem(64, 8)
    r = runner(shape, ax, seeds, 1e10, 400, [0, 200, 400], 100, "fp32")
    ens, fin, st = r.forward(THETA, record_gradient=False)
    ens2, fin2, _ = r.forward(THETA, record_gradient=True)
    assert np.array_equal(ens, ens2) and fin == fin2
    r.forward(THETA, record_gradient=False)
    with pytest.raises(ValueError):
        r.reverse([0.0, 1.0, 1.0], 0.0)


def test_particle_gradients_split_the_total():
    print("CASE adjoint_particle_split", flush=True)
    shape, ax, seeds = problem(96, 9)
    ks, N = [0, 150, 300], 300
    r = runner(shape, ax, seeds, 1e10, N, ks, 100, "fp64")
    r.forward(THETA)
    w = np.array([0.0, 2.0, -1.0])
    total = r.reverse(w, 0.25)
    per = r.particle_gradients()
    assert np.allclose(per.sum(0), total, rtol=1e-12, atol=0)
    # one particle's part = the oracle tangent of that particle alone
    p = 17
    q, dq = core.adjoint_reduced(THETA, shape[p:p + 1], MS, T, DT)
    pr = O.path_problem(ax[p:p + 1], ax[p:p + 1], q[:, 0], q[:, 1], q[:, 2], q[:, 3], 1e10 * DT,
                        seeds[p:p + 1], N, ks)
    tan, tan_f = O.path_tangent(pr, np.ascontiguousarray(np.transpose(dq, (0, 2, 1))))
    ref = (MS / 96) * (w @ tan) + (0.25 / 96) * tan_f
    assert np.all(np.abs(per[p] - ref) <= 1e-9 * np.abs(ref)), (per[p], ref)
