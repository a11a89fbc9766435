"""The FLOP counts used by the rooflines come from the oracle's arithmetic.

oracle/flop_count.py runs the step formulas on a counting number type.  It is
pinned two ways: (1) the counted Heun and implicit-midpoint steps reproduce,
value for value, what the compiled oracle's orc_heun and orc_implicit_midpoint
(restating lib/integrators.cpp:372-405, 576-651 and
lib/optimisation.cpp:81-149) give for one particle fed the oracle's own
field, drift, diffusion and Jacobian functions, with the same Newton
iteration count; (2) the Heun count equals the 278 FLOP per particle-step
derived by hand from the reference's formulas (field 13, drift 42, diffusion
39 per stage, predictor 33, corrector 57), so the counting convention is
the one the headline roofline uses.  (3) The counted Cash-Karp step of the
discrete orientation model reproduces orc_rk45 (lib/integrators.cpp:152-251)
on the compiled oracle's master equation value for value, rejected attempts
included.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import flop_count as F
from oracle import oracle as O

L = O.lib()
ptr = O.ptr


def one_particle(seed):
    rng = np.random.default_rng(seed)
    m = rng.normal(size=3)
    m /= np.linalg.norm(m)
    e = rng.normal(size=3)
    e /= np.linalg.norm(e)
    return m, e, 0.9 + 0.2 * rng.random(), 0.1, 0.05 + 0.02 * rng.random(), rng.normal(size=3)


def oracle_sde(e, k, al, sr, happ, jac):
    """field + drift + diffusion (+ Jacobians) through the compiled oracle"""
    ev, kv = O.f64(e), O.f64([k])
    alv, srv = O.f64([al]), O.f64([sr])

    def parts(state):
        h = np.zeros(3)
        L.orc_field_multi_add_uniaxial(ptr(h), ptr(state), ptr(ev), ptr(kv), 1)
        L.orc_field_multi_add_applied_z(ptr(h), happ, 1)
        a = np.zeros(3)
        L.orc_llg_drift(ptr(a), ptr(state), al, ptr(h))
        b = np.zeros(9)
        L.orc_llg_multi_diffusion(ptr(b), ptr(state), ptr(srv), ptr(alv), 1)
        if not jac:
            return a, b
        hj = np.zeros(9)
        L.orc_field_multi_add_uniaxial_jacobian(ptr(hj), ptr(ev), ptr(kv), 1)
        ad = np.zeros(9)
        L.orc_llg_multi_drift_quasijacobian(ptr(ad), ptr(state), ptr(alv), ptr(h), ptr(hj), 1)
        bd = np.zeros(27)
        L.orc_llg_multi_diffusion_jacobian(ptr(bd), ptr(state), ptr(srv), ptr(alv), 1)
        return a, b, ad, bd
    return parts


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_counted_heun_step_is_the_oracle_step(seed):
    m, e, k, al, sr, w = one_particle(seed)
    dt, happ = 2e-4, 0.13
    parts = oracle_sde(e, k, al, sr, happ, jac=False)
    nxt, dr, tdr, df, tdf = (np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(9), np.zeros(9))

    def cb(ctx, drift, diff, state, tt):
        a, b = parts(np.ctypeslib.as_array(state, (3,)).copy())
        for i in range(3):
            drift[i] = a[i]
        for i in range(9):
            diff[i] = b[i]

    fn = O.SDE_FN(cb)
    L.orc_heun(ptr(nxt), ptr(dr), ptr(tdr), ptr(df), ptr(tdf), ptr(O.f64(m)), ptr(O.f64(w)), fn,
               None, 3, 3, 0.0, dt)
    counted, flops = F.heun_step(m, e, k, al, sr, happ, happ, dt, w)
    assert np.allclose(counted, nxt, rtol=0, atol=1e-15)
    assert flops == 278


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_counted_implicit_step_is_the_oracle_step(seed):
    m, e, k, al, sr, w = one_particle(seed)
    dt, happ, eps = 1e-2, 0.13, 1e-9
    parts = oracle_sde(e, k, al, sr, happ, jac=True)
    x, dwm, aw, bw = np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(9)
    ad, bd = np.zeros(9), np.zeros(27)
    xg, xt, xj = np.zeros(3), np.zeros(3), np.zeros(9)
    ipiv = np.zeros(3, dtype=np.int32)
    iters = C.c_int(0)

    def cb(ctx, a, b, adash, bdash, state, ta, tb):
        A, B, AD, BD = parts(np.ctypeslib.as_array(state, (3,)).copy())
        for i in range(3):
            a[i] = A[i]
        for i in range(9):
            b[i] = B[i]
            adash[i] = AD[i]
        for i in range(27):
            bdash[i] = BD[i]

    fn = O.SDEJAC_FN(cb)
    code = L.orc_implicit_midpoint(ptr(x), ptr(dwm), ptr(aw), ptr(bw), ptr(ad), ptr(bd),
                                   ptr(xg), ptr(xt), ptr(xj), ptr(ipiv), ptr(O.f64(m)),
                                   ptr(O.f64(w)), fn, None, 3, 3, 0.0, dt, eps, 1000,
                                   C.byref(iters))
    assert code == 0
    counted, n_it, other, per_iter = F.implicit_step(m, e, k, al, sr, happ, happ, dt, w, eps)
    assert n_it == iters.value
    assert np.allclose(counted, x, rtol=0, atol=1e-13)
    per_iter_ref, other_ref = F.implicit_flops()
    assert (per_iter, other) == (per_iter_ref, other_ref)


def test_counting_sees_each_operation():
    F.reset()
    a, b = F.Count(1.5), F.Count(2.0)
    c = a * b + 3.0 - b / 4.0      # 4 operations
    _ = -c                          # negation is not counted
    _ = F.csqrt(c)
    assert F.Count.flops == 4 and F.Count.sqrts == 1
    assert math.isclose(float(c), 1.5 * 2.0 + 3.0 - 0.5)


def oracle_dom_ode(prm):
    """the master equation through the compiled oracle's field, transition
    matrix and matrix-vector product (lib/dom.cpp:30-93)"""
    k, v, T, ms, al, h, f = (float(x) for x in prm)

    def ode(x, t):
        W = np.zeros(4)
        L.orc_dom_transition_matrix(ptr(W), k, v, T, L.orc_field_sinusoidal(t, h, f), ms, al)
        dx = np.zeros(2)
        L.orc_master_equation(ptr(dx), ptr(W), ptr(O.f64(x)), 2)
        return dx
    return ode


@pytest.mark.parametrize("h,tol", [(1e-7, 1e-9), (1e-5, 1e-9), (1e-3, 1e-12), (3e-8, 1e-10)])
def test_counted_rk45_step_is_the_oracle_step(h, tol):
    prm = F.dom_params()
    x = [0.7, 0.3]
    t = 1e-7
    nxt, _, t_new, h_new = O.rk45(oracle_dom_ode(prm), x, t, h, tol)
    F.reset()
    cn, ct, ch, acc, rej = F.rk45_step(lambda xx, tt: F.dom_derivs(prm, xx, tt),
                                       [F.Count(v) for v in x], F.Count(t), F.Count(h),
                                       F.Count(tol))
    assert np.array_equal([float(v) for v in cn], nxt)
    assert float(ct) == t_new and float(ch) == h_new
    # each attempt costs six derivative evaluations of 44 FLOP and the
    # stage and error arithmetic around them
    assert acc == 401 and all(r == 400 for r in rej)


def test_bench_constants_are_the_counts():
    import bench
    assert bench.FLOP_PER_PARTICLE_STEP == F.heun_flops()
    assert (bench.FLOP_IMPLICIT_PER_NEWTON, bench.FLOP_IMPLICIT_PER_STEP) == F.implicit_flops()
    assert bench.FLOP_DOM_PER_STEP == F.dom_flops()[0]
