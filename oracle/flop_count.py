"""Algorithmic FLOP counts of the reference's step arithmetic (test
infrastructure: used by bench.py's rooflines and checked by the tests, never
by the product path).

The counts are taken by running the oracle's formulas (oracle/sllg_oracle.c,
which restates lib/llg.cpp, lib/field.cpp, lib/integrators.cpp and
lib/optimisation.cpp function by function) on a counting number type for one
particle (N = 1, so every (3N)-sized loop of the reference has its N = 1
trip count).  Convention: each add, sub, mul and div is one FLOP; negation,
comparisons, copies, the per-step sin of the applied field and the noise
generation are not counted; sqrt(dt) is hoisted (counted once per step,
never); square roots of norms are counted separately as `sqrt`.  The same
convention gives the Heun step's 278 FLOP per particle-step, which the tests
check.

  heun_flops()              lib/integrators.cpp:372-405 with the sde of
                            lib/simulation.cpp:328-336
  implicit_flops()          lib/integrators.cpp:576-651 with
                            optimisation::newton_raphson_noinv
                            (lib/optimisation.cpp:81-149) and the sde of
                            lib/simulation.cpp:306-325: FLOP per Newton
                            iteration and per step outside the iterations
  dom_flops()               one accepted Cash-Karp step of integrator::rk45
                            (lib/integrators.cpp:152-251) on the master
                            equation of lib/dom.cpp:30-93 with the sine field
                            of lib/field.cpp:37-40: FLOP per accepted attempt
                            and per rejected attempt

For the discrete orientation model the transcendental calls (sin, exp and
the pow(sigma, 1.5), pow(err, -0.2) of the rates and the step control) are
not counted, pow(x, 2) counts as one multiply and pow(x, 0.5) of a norm as a
sqrt; each derivative evaluation counts the particle constants (sigma,
tau_N) that the reference recomputes there, as written.
"""
import math


class Count:
    """A float that counts the arithmetic done on it (shared counter)."""
    flops = 0
    sqrts = 0

    __slots__ = ("v",)

    def __init__(self, v):
        self.v = float(v.v if isinstance(v, Count) else v)

    @staticmethod
    def _val(o):
        return o.v if isinstance(o, Count) else float(o)

    def _op(self, o, f):
        Count.flops += 1
        return Count(f(self.v, Count._val(o)))

    def __add__(self, o): return self._op(o, lambda a, b: a + b)
    def __radd__(self, o): return self._op(o, lambda a, b: b + a)
    def __sub__(self, o): return self._op(o, lambda a, b: a - b)
    def __rsub__(self, o): return self._op(o, lambda a, b: b - a)
    def __mul__(self, o): return self._op(o, lambda a, b: a * b)
    def __rmul__(self, o): return self._op(o, lambda a, b: b * a)
    def __truediv__(self, o): return self._op(o, lambda a, b: a / b)
    def __rtruediv__(self, o): return self._op(o, lambda a, b: b / a)
    def __neg__(self): return Count(-self.v)
    def __abs__(self): return Count(abs(self.v))
    def __gt__(self, o): return self.v > Count._val(o)
    def __lt__(self, o): return self.v < Count._val(o)
    def __ge__(self, o): return self.v >= Count._val(o)
    def __ne__(self, o): return self.v != Count._val(o)
    def __eq__(self, o): return self.v == Count._val(o)
    __hash__ = None

    def __float__(self): return self.v


def csqrt(x):
    Count.sqrts += 1
    return Count(math.sqrt(float(x)))


def reset():
    Count.flops = 0
    Count.sqrts = 0


# ---------------------------------------------------------------- formulas
# each follows the oracle function of the same name (oracle/sllg_oracle.c)

def drift(m, al, h):
    """orc_llg_drift, lib/llg.cpp:14-29"""
    return [m[2] * h[1] - m[1] * h[2]
            + al * (h[0] * (m[1] * m[1] + m[2] * m[2]) - m[0] * (m[1] * h[1] + m[2] * h[2])),
            m[0] * h[2] - m[2] * h[0]
            + al * (h[1] * (m[0] * m[0] + m[2] * m[2]) - m[1] * (m[0] * h[0] + m[2] * h[2])),
            m[1] * h[0] - m[0] * h[1]
            + al * (h[2] * (m[0] * m[0] + m[1] * m[1]) - m[2] * (m[0] * h[0] + m[1] * h[1]))]


def diffusion(m, sr, al):
    """orc_llg_diffusion, lib/llg.cpp:92-106 (row-major 3x3)"""
    return [al * sr * (m[1] * m[1] + m[2] * m[2]),
            sr * (m[2] - al * m[0] * m[1]),
            -sr * (m[1] + al * m[0] * m[2]),
            -sr * (m[2] + al * m[0] * m[1]),
            al * sr * (m[0] * m[0] + m[2] * m[2]),
            sr * (m[0] - al * m[1] * m[2]),
            sr * (m[1] - al * m[0] * m[2]),
            -sr * (m[0] + al * m[1] * m[2]),
            al * sr * (m[0] * m[0] + m[1] * m[1])]


def drift_jacobian(m, a, h, hj):
    """orc_llg_drift_jacobian, lib/llg.cpp:68-81"""
    return [m[2] * hj[3] - m[1] * hj[6]
            + a * (-m[1] * h[1] - m[2] * h[2] + (m[1] * m[1] + m[2] * m[2]) * hj[0]
                   - m[0] * (m[1] * hj[3] + m[2] * hj[6])),
            -h[2] + m[2] * hj[4] - m[1] * hj[7]
            + a * (2 * m[1] * h[0] + (m[1] * m[1] + m[2] * m[2]) * hj[1]
                   - m[0] * (h[1] + m[1] * hj[4] + m[2] * hj[7])),
            h[1] + m[2] * hj[5] - m[1] * hj[8]
            + a * (2 * m[2] * h[0] + (m[1] * m[1] + m[2] * m[2]) * hj[2]
                   - m[0] * (h[2] + m[1] * hj[5] + m[2] * hj[8])),
            h[2] - m[2] * hj[0] + m[0] * hj[6]
            + a * (2 * m[0] * h[1] + (m[0] * m[0] + m[2] * m[2]) * hj[3]
                   - m[1] * (h[0] + m[0] * hj[0] + m[2] * hj[6])),
            -m[2] * hj[1] + m[0] * hj[7]
            + a * (-m[0] * h[0] - m[2] * h[2] + (m[0] * m[0] + m[2] * m[2]) * hj[4]
                   - m[1] * (m[0] * hj[1] + m[2] * hj[7])),
            -h[0] - m[2] * hj[2] + m[0] * hj[8]
            + a * (2 * m[2] * h[1] + (m[0] * m[0] + m[2] * m[2]) * hj[5]
                   - m[1] * (h[2] + m[0] * hj[2] + m[2] * hj[8])),
            -h[1] + m[1] * hj[0] - m[0] * hj[3]
            + a * (2 * m[0] * h[2] + (m[0] * m[0] + m[1] * m[1]) * hj[6]
                   - m[2] * (h[0] + m[0] * hj[0] + m[1] * hj[3])),
            h[0] + m[1] * hj[1] - m[0] * hj[4]
            + a * (2 * m[1] * h[2] + (m[0] * m[0] + m[1] * m[1]) * hj[7]
                   - m[2] * (h[1] + m[0] * hj[1] + m[1] * hj[4])),
            m[1] * hj[2] - m[0] * hj[5]
            + a * (-m[0] * h[0] - m[1] * h[1] + (m[0] * m[0] + m[1] * m[1]) * hj[8]
                   - m[2] * (m[0] * hj[2] + m[1] * hj[5]))]


def diffusion_jacobian(m, sr, al):
    """orc_llg_diffusion_jacobian, lib/llg.cpp:118-158: jc[z + 3y + 9x]"""
    z = Count(0.0)
    return [z, 2 * al * sr * m[1], 2 * al * sr * m[2],
            -al * sr * m[1], -al * sr * m[2], Count(sr),
            -al * sr * m[2], -Count(sr), -al * sr * m[0],
            -al * sr * m[1], -al * sr * m[0], -Count(sr),
            2 * al * sr * m[0], z, 2 * al * sr * m[2],
            Count(sr), -al * sr * m[2], -al * sr * m[1],
            -al * sr * m[2], Count(sr), -al * sr * m[0],
            -Count(sr), -al * sr * m[2], -al * sr * m[1],
            2 * al * sr * m[0], 2 * al * sr * m[2], z]


def field(m, e, k, happ_sin):
    """orc_heff for N = 1: zero, orc_field_multi_add_uniaxial (lib/field.cpp:
    116-132), orc_field_multi_add_applied_z (:81-88) with the applied value
    h sin(2 pi f t) given (its sin is not counted)"""
    h = [Count(0.0), Count(0.0), Count(0.0)]
    dot = m[0] * e[0] + m[1] * e[1] + m[2] * e[2]
    dot = dot * k
    h = [h[0] + dot * e[0], h[1] + dot * e[1], h[2] + dot * e[2]]
    h[2] = h[2] + happ_sin
    return h


def field_jacobian(e, k):
    """orc_field_multi_add_uniaxial_jacobian (lib/field.cpp:159-174) for
    N = 1: jac[i][j] = k e_i e_j"""
    return [k * e[i] * e[j] for i in range(3) for j in range(3)]


def dnrm2(x):
    """orc_dnrm2: the plain 2-norm"""
    s = Count(0.0)
    for v in x:
        s = s + v * v
    return csqrt(s)


def dgesv(a, b):
    """orc_dgesv for n = 3: partial pivoting, reciprocal scaling, solve"""
    n = 3
    a = list(a)
    b = list(b)
    piv = []
    for j in range(n):
        p = j
        best = abs(a[j * n + j])
        for i in range(j + 1, n):
            if abs(a[i * n + j]) > best:
                best = abs(a[i * n + j])
                p = i
        piv.append(p)
        if p != j:
            for k in range(n):
                a[j * n + k], a[p * n + k] = a[p * n + k], a[j * n + k]
        rcp = 1.0 / a[j * n + j]
        for i in range(j + 1, n):
            a[i * n + j] = a[i * n + j] * rcp
        for i in range(j + 1, n):
            l = a[i * n + j]
            for k in range(j + 1, n):
                a[i * n + k] = a[i * n + k] - l * a[j * n + k]
    for j in range(n):
        if piv[j] != j:
            b[j], b[piv[j]] = b[piv[j]], b[j]
    for i in range(n):
        for k in range(i):
            b[i] = b[i] - a[i * n + k] * b[k]
    for i in range(n - 1, -1, -1):
        for k in range(i + 1, n):
            b[i] = b[i] - a[i * n + k] * b[k]
        b[i] = b[i] / a[i * n + i]
    return b


def sde_jac(m, e, k, al, sr, happ_sin):
    """orc_sde_jac for N = 1 (lib/simulation.cpp:306-325)"""
    h = field(m, e, k, happ_sin)
    return (drift(m, al, h), diffusion(m, sr, al),
            drift_jacobian(m, al, h, field_jacobian(e, k)), diffusion_jacobian(m, sr, al))


# ---------------------------------------------------------------- steps

def _params(seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    m = rng.normal(size=3)
    m /= np.linalg.norm(m)
    e = rng.normal(size=3)
    e /= np.linalg.norm(e)
    return (m, e, 0.9 + 0.2 * rng.random(), 0.1, 0.05 + 0.02 * rng.random(),
            0.17 * rng.random(), 2e-4, rng.normal(size=3))


def heun_step(m, e, k, al, sr, hs1, hs2, dt, w):
    """orc_heun with orc_sde_nojac for N = 1; returns (next, flops)"""
    reset()
    m = [Count(v) for v in m]
    e = [Count(v) for v in e]
    w = [Count(v) for v in w]
    k, al, sr, hs1, hs2 = Count(k), Count(al), Count(sr), Count(hs1), Count(hs2)
    sdt = Count(math.sqrt(dt))        # hoisted
    dt = Count(dt)
    a = drift(m, al, field(m, e, k, hs1))
    b = diffusion(m, sr, al)
    nxt = []
    for i in range(3):
        v = m[i] + dt * a[i]
        for j in range(3):
            v = v + b[j + 3 * i] * w[j] * sdt
        nxt.append(v)
    ta = drift(nxt, al, field(nxt, e, k, hs2))
    tb = diffusion(nxt, sr, al)
    out = []
    for i in range(3):
        v = m[i] + 0.5 * dt * (ta[i] + a[i])
        for j in range(3):
            v = v + 0.5 * w[j] * sdt * (tb[j + 3 * i] + b[j + 3 * i])
        out.append(v)
    return [float(v) for v in out], Count.flops


def heun_flops():
    m, e, k, al, sr, hs, dt, w = _params(1)
    return heun_step(m, e, k, al, sr, hs, hs, dt, w)[1]


def implicit_step(x0, e, k, al, sr, hs_a, hs_mid, dt, w, eps=1e-9, max_iter=1000):
    """orc_implicit_midpoint for N = 1 (lib/integrators.cpp:576-651, the FJ
    closure :617-637, newton_raphson_noinv lib/optimisation.cpp:81-149).
    hs_a: applied field at t (the guess), hs_mid: at t + dt/2.  Returns
    (x, iterations, flops outside the iterations, flops per iteration)."""
    reset()
    x0 = [Count(v) for v in x0]
    e = [Count(v) for v in e]
    k, al, sr, hs_a, hs_mid = Count(k), Count(al), Count(sr), Count(hs_a), Count(hs_mid)
    ah = math.sqrt(2 * 1000 * abs(math.log(dt)))   # per-run constants
    sdt = Count(math.sqrt(dt))
    dt = Count(dt)
    dwm = [Count(max(-ah, min(ah, wj))) * sdt for wj in w]
    a, b, _, _ = sde_jac(x0, e, k, al, sr, hs_a)
    guess = []
    for i in range(3):
        v = a[i] * dt
        for j in range(3):
            v = v + b[i * 3 + j] * dwm[j]
        guess.append(v)
    guess = [(guess[i] + x0[i]) / 2 for i in range(3)]
    x = list(guess)
    tol = eps * dnrm2(x)
    pre = Count.flops
    err = 2 * float(tol)
    iters = 0
    per_iter = None
    while err > float(tol) and iters < max_iter:
        start = Count.flops
        xt = list(x)
        a, b, ad, bd = sde_jac(x, e, k, al, sr, hs_mid)
        f = []
        for i in range(3):
            v = x[i] - 0.5 * a[i] * dt - x0[i]
            for j in range(3):
                v = v - 0.5 * b[i * 3 + j] * dwm[j]
            f.append(v)
        jac = []
        for i in range(3):
            for j in range(3):
                v = Count(1.0 if i == j else 0.0) - 0.5 * ad[i * 3 + j]
                for kk in range(3):
                    v = v - 0.5 * bd[i * 9 + kk * 3 + j] * dwm[kk]
                jac.append(v)
        x = [v * -1 for v in f]
        x = dgesv(jac, x)
        err = float(dnrm2(x))
        x = [x[i] + xt[i] for i in range(3)]
        iters += 1
        n = Count.flops - start
        assert per_iter is None or per_iter == n
        per_iter = n
    before_end = Count.flops
    out = [2 * x[i] - x0[i] for i in range(3)]
    post = Count.flops - before_end
    return [float(v) for v in out], iters, pre + post, per_iter


def implicit_flops():
    """(FLOP per Newton iteration, FLOP per step outside the iterations)"""
    m, e, k, al, sr, hs, dt, w = _params(2)
    _, iters, other, per_iter = implicit_step(m, e, k, al, sr, hs, hs, 1e-2, w)
    assert iters >= 1
    return per_iter, other


if __name__ == "__main__":
    print("heun FLOP per particle-step", heun_flops())
    pi, po = implicit_flops()
    print("implicit FLOP per Newton iteration", pi, "per step besides", po)


# ---------------------------------------------------------------- discrete orientation model
KB = 1.38064852e-23      # include/constants.hpp:10
GAMMA = 1.76086e11       # include/constants.hpp:12


def cpow(x, e):
    """pow as the step arithmetic uses it: a square is one multiply, a square
    root of a norm a sqrt, any other exponent a math function (not counted)."""
    if e == 2:
        return x * x
    if e == 0.5:
        return csqrt(x)
    return Count(math.pow(float(x), e))


def field_sinusoidal(t, h, f):
    """orc_field_sinusoidal (lib/field.cpp:37-40): h sin(2 pi f t)."""
    return h * Count(math.sin(float(Count(2 * math.pi) * f * t)))


def dom_transition_matrix(k, v, T, h, ms, alpha):
    """orc_dom_transition_matrix (lib/dom.cpp:30-56)."""
    sigma = k * v / KB / T
    taun = v * ms * (1 + alpha * alpha) / 2.0 / GAMMA / alpha / KB / T
    e1 = sigma * (1 - h) * (1 - h)
    e2 = sigma * (1 + h) * (1 + h)
    pre = taun * math.sqrt(math.pi) / cpow(sigma, 1.5) / (1 - h * h)
    r1 = 1.0 / pre * (1 - h) * Count(math.exp(-float(e1)))
    r2 = 1.0 / pre * (1 + h) * Count(math.exp(-float(e2)))
    return [-r2, r1, r2, -r1]


def master_equation(W, x):
    """orc_master_equation (lib/dom.cpp:59-77), dim 2."""
    out = []
    for i in range(2):
        acc = Count(0.0)
        for j in range(2):
            acc = acc + W[i * 2 + j] * x[j]
        out.append(acc)
    return out


def dom_derivs(prm, x, t):
    """orc_dom_master (lib/dom.cpp:80-93) with the sine field."""
    k, v, T, ms, alpha, h, f = prm
    return master_equation(dom_transition_matrix(k, v, T, field_sinusoidal(t, h, f), ms, alpha),
                           x)


CK = dict(C11=0.2, C21=3.0 / 40.0, C22=9.0 / 40.0, C31=3.0 / 10.0, C32=-9.0 / 10.0, C33=6.0 / 5.0,
          C41=-11.0 / 54.0, C42=2.5, C43=-70.0 / 27.0, C44=35.0 / 27.0, C51=1631.0 / 55296.0,
          C52=175.0 / 512.0, C53=575.0 / 13824.0, C54=44275.0 / 110592.0, C55=253.0 / 4096.0,
          H1=0.2, H2=0.3, H3=0.6, H4=1.0, H5=7.0 / 8.0, X11=37.0 / 378.0, X13=250.0 / 621.0,
          X14=125.0 / 594.0, X16=512.0 / 1771.0, X21=2825.0 / 27648.0, X23=18575.0 / 48384.0,
          X24=13525.0 / 55296.0, X25=277.0 / 14336.0, X26=0.25)


def rk45_step(derivs, x, t, h, tol):
    """orc_rk45 (lib/integrators.cpp:152-251, Cash-Karp tableau of
    include/integrators.hpp:128-160) on Count numbers, n = 2.  Returns
    (next_state, t_new, h_new, flops of the accepted attempt, flops of each
    rejected attempt)."""
    c = CK
    n = len(x)
    rejected = []
    while True:
        f0 = Count.flops
        k1 = derivs(x, t)
        y = [k1[i] * h * c["C11"] + x[i] for i in range(n)]
        k2 = derivs(y, t + h * c["H1"])
        y = [x[i] + h * (c["C21"] * k1[i] + c["C22"] * k2[i]) for i in range(n)]
        k3 = derivs(y, t + h * c["H2"])
        y = [x[i] + h * (c["C31"] * k1[i] + c["C32"] * k2[i] + c["C33"] * k3[i])
             for i in range(n)]
        k4 = derivs(y, t + h * c["H3"])
        y = [x[i] + h * (c["C41"] * k1[i] + c["C42"] * k2[i] + c["C43"] * k3[i] +
                         c["C44"] * k4[i]) for i in range(n)]
        k5 = derivs(y, t + h * c["H4"])
        y = [x[i] + h * (c["C51"] * k1[i] + c["C52"] * k2[i] + c["C53"] * k3[i] +
                         c["C54"] * k4[i] + c["C55"] * k5[i]) for i in range(n)]
        k6 = derivs(y, t + h * c["H5"])
        tmp = [x[i] + h * (c["X11"] * k1[i] + c["X13"] * k3[i] + c["X14"] * k4[i] +
                           c["X16"] * k6[i]) for i in range(n)]
        nxt = [x[i] + h * (c["X21"] * k1[i] + c["X23"] * k3[i] + c["X24"] * k4[i] +
                           c["X25"] * k5[i] + c["X26"] * k6[i]) for i in range(n)]
        err = Count(0.0)
        mag = Count(0.0)
        for i in range(n):
            mag = mag + x[i] * x[i]
        mag = cpow(mag, 0.5)
        for i in range(n):
            err = err + cpow(abs(tmp[i] - nxt[i]), 2)
        err = cpow(err, 0.5)
        err = err / (Count(float(n)) * tol * (1 + mag))
        if err < 1.0:
            break
        f = 0.84 * cpow(err, -0.2)
        f = Count(0.1) if abs(f) < 0.1 else f
        h = h * f
        rejected.append(Count.flops - f0)
    t_new = t + h
    f = Count(5.0) if err == 0.0 else 0.84 * cpow(err, -0.2)
    f = Count(5.0) if f > 5 else f
    h_new = f * h
    return nxt, t_new, h_new, Count.flops - f0, rejected


def dom_params():
    """A 10 nm particle of the AC-field workload: K = 23 kJ/m^3, Ms = 400 kA/m,
    alpha = 0.1, T = 300 K, reduced field amplitude 0.174, 300 kHz."""
    v = 4.0 / 3.0 * math.pi * (10e-9) ** 3
    return (Count(23e3), Count(v), Count(300.0), Count(4e5), Count(0.1), Count(0.174),
            Count(3e5))


def dom_flops():
    """(FLOP of an accepted rk45 attempt, FLOP of a rejected one) on the
    master equation, counted as the oracle's arithmetic is written."""
    prm = dom_params()
    x = [Count(0.7), Count(0.3)]
    reset()
    # accepted at once with a non-zero error estimate (the usual case; an
    # exactly zero estimate skips one multiply)
    _, _, _, acc, rej = rk45_step(lambda xx, tt: dom_derivs(prm, xx, tt), x, Count(1e-7),
                                  Count(1e-7), Count(1e-9))
    assert not rej
    reset()
    # a step far too long for the tolerance is rejected first
    _, _, _, _, rej = rk45_step(lambda xx, tt: dom_derivs(prm, xx, tt), x, Count(1e-7),
                                Count(1e-3), Count(1e-12))
    return acc, rej[0]
