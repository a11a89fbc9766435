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
the one the headline roofline uses.
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


def test_bench_constants_are_the_counts():
    import bench
    assert bench.FLOP_PER_PARTICLE_STEP == F.heun_flops()
    assert (bench.FLOP_IMPLICIT_PER_NEWTON, bench.FLOP_IMPLICIT_PER_STEP) == F.implicit_flops()
