"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package magpy_amd.

The C library restates owlas/magpy's FP64 sLLG path (lib/llg.cpp,
lib/field.cpp, lib/integrators.cpp, lib/optimisation.cpp, lib/simulation.cpp,
lib/rng.cpp, lib/trap.cpp); see sllg_oracle.c for per-function citations.
`ref()` loads oracle/_ref/libmagpy_ref.so, the reference's own building
blocks compiled from source, where it was built.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libmagpy_ref.so")
REFERENCE_TREE = "/root/reference"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)

FIELD = {"sine": 0, "square": 1, "constant": 2}
RNG_MT, RNG_ARRAY, RNG_PHILOX = 0, 1, 2

SDE_FN = C.CFUNCTYPE(None, C.c_void_p, _dp, _dp, _dp, C.c_double)
SDEJAC_FN = C.CFUNCTYPE(None, C.c_void_p, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double)
FJ_FN = C.CFUNCTYPE(None, C.c_void_p, _dp, _dp, _dp)
ODE_FN = C.CFUNCTYPE(None, C.c_void_p, _dp, _dp, C.c_double)
FIELD_SQUARE_F = 3


class PathProblem(C.Structure):
    """orc_path_problem"""
    _fields_ = [("n_particles", C.c_size_t), ("axis", _dp), ("m0", _dp), ("dt", _dp),
                ("sig", _dp), ("alpha", _dp), ("happ", _dp), ("fdt", C.c_double),
                ("seed", C.POINTER(C.c_uint64)), ("tag", C.c_uint32), ("n_steps", C.c_int64),
                ("sample_steps", _i64p), ("n_samples", C.c_size_t)]


class Stats(C.Structure):
    _fields_ = [("n_steps", C.c_int64), ("newton_iters", C.c_int64),
                ("newton_failures", C.c_int64), ("rng_exhausted", C.c_int)]


def build(ref=True):
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir(os.path.join(REFERENCE_TREE, "lib")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", "REF=" + REFERENCE_TREE], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_u01.restype = C.c_double
        L.orc_u01.argtypes = [C.c_uint32]
        L.orc_wiener_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64,
                                         C.c_size_t, _dp]
        for nm, args in [
            ("orc_llg_drift", [_dp, _dp, C.c_double, _dp]),
            ("orc_llg_diffusion", [_dp, _dp, C.c_double, C.c_double]),
            ("orc_llg_drift_jacobian", [_dp, _dp, C.c_double, _dp, _dp]),
            ("orc_llg_diffusion_jacobian", [_dp, _dp, C.c_double, C.c_double]),
            ("orc_llg_multi_diffusion", [_dp, _dp, _dp, _dp, C.c_size_t]),
            ("orc_llg_multi_drift_quasijacobian", [_dp, _dp, _dp, _dp, _dp, C.c_size_t]),
            ("orc_llg_multi_diffusion_jacobian", [_dp, _dp, _dp, _dp, C.c_size_t]),
            ("orc_field_uniaxial", [_dp, _dp, _dp]),
            ("orc_field_uniaxial_jacobian", [_dp, _dp]),
            ("orc_field_multi_add_uniaxial", [_dp, _dp, _dp, _dp, C.c_size_t]),
            ("orc_field_multi_add_uniaxial_jacobian", [_dp, _dp, _dp, C.c_size_t]),
            ("orc_field_multi_add_applied_z", [_dp, C.c_double, C.c_size_t]),
            ("orc_dipolar_add_p2p_term", [_dp, C.c_double, C.c_double, _dp, _dp, C.c_double]),
            ("orc_field_multi_add_dipolar", [_dp, C.c_double, C.c_double, _dp, _dp, _dp, _dp,
                                             C.c_size_t]),
            ("orc_pair_distance_vectors", [_dp, _dp, C.c_size_t]),
            ("orc_pair_distance_magnitudes", [_dp, _dp, C.c_size_t]),
            ("orc_pair_unit_vectors", [_dp, _dp, _dp, C.c_size_t]),
            ("orc_heun", [_dp, _dp, _dp, _dp, _dp, _dp, _dp, SDE_FN, C.c_void_p, C.c_size_t,
                          C.c_size_t, C.c_double, C.c_double]),
        ]:
            getattr(L, nm).argtypes = args
            getattr(L, nm).restype = None
        for nm in ("orc_field_sinusoidal", "orc_field_square"):
            getattr(L, nm).restype = C.c_double
            getattr(L, nm).argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_dipolar_prefactor.restype = C.c_double
        L.orc_dipolar_prefactor.argtypes = [C.c_double, C.c_double]
        L.orc_trapezoidal.restype = C.c_double
        L.orc_trapezoidal.argtypes = [_dp, _dp, C.c_size_t]
        L.orc_dgesv.restype = C.c_int
        L.orc_dgesv.argtypes = [C.c_int, _dp, _ip, _dp]
        L.orc_newton_noinv.restype = C.c_int
        L.orc_newton_noinv.argtypes = [_dp, _dp, _dp, _ip, _ip, FJ_FN, C.c_void_p, _dp, C.c_int,
                                       C.c_double, C.c_size_t]
        L.orc_implicit_midpoint.restype = C.c_int
        L.orc_implicit_midpoint.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip,
                                            _dp, _dp, SDEJAC_FN, C.c_void_p, C.c_size_t,
                                            C.c_size_t, C.c_double, C.c_double, C.c_double,
                                            C.c_size_t, _ip]
        L.orc_path_phase.restype = C.c_double
        L.orc_path_phase.argtypes = [C.c_double, C.c_int64]
        L.orc_path_step.restype = None
        L.orc_path_step.argtypes = [_dp, _dp, _dp] + [C.c_double] * 6
        L.orc_path_run.restype = None
        L.orc_path_run.argtypes = [C.POINTER(PathProblem), _dp, _dp, _dp]
        L.orc_path_tangent.restype = None
        L.orc_path_tangent.argtypes = [C.POINTER(PathProblem), C.c_size_t, _dp, _dp, _dp]
        L.orc_field_square_fourier.restype = C.c_double
        L.orc_field_square_fourier.argtypes = [C.c_double, C.c_double, C.c_double, C.c_size_t]
        L.orc_dom_transition_matrix.restype = None
        L.orc_dom_transition_matrix.argtypes = [_dp] + [C.c_double] * 6
        L.orc_master_equation.restype = None
        L.orc_master_equation.argtypes = [_dp, _dp, _dp, C.c_size_t]
        L.orc_rk45.restype = None
        L.orc_rk45.argtypes = [_dp] * 10 + [_dp, ODE_FN, C.c_void_p, C.c_size_t, C.c_double]
        L.orc_dom_ensemble_dynamics.restype = None
        L.orc_dom_ensemble_dynamics.argtypes = [C.c_double] * 5 + [C.c_int, C.c_double,
                                                                  C.c_double, C.c_size_t, _dp,
                                                                  C.c_double, C.c_double,
                                                                  C.c_size_t, _dp, _dp, _dp]
        L.orc_dnrm2.restype = C.c_double
        L.orc_dnrm2.argtypes = [C.c_int, _dp]
        L.orc_sample_schedule.restype = C.c_int64
        L.orc_sample_schedule.argtypes = [C.c_double, C.c_double, C.c_size_t, _i64p]
        L.orc_full_dynamics.restype = C.c_int
        L.orc_full_dynamics.argtypes = [
            _dp, _dp, _dp, _dp, _dp, C.c_size_t, C.c_double, C.c_double, C.c_double, C.c_int,
            C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_size_t, C.c_int64,
            C.c_int, C.c_double, C.c_double, C.c_int, _dp, C.c_size_t, _dp, _dp, _dp, _dp, _dp,
            C.POINTER(Stats)]
        _lib = L
    return _lib


def ref():
    """The reference's building blocks compiled from source, or None."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        for nm in ("ref_field_sinusoidal", "ref_field_square"):
            getattr(R, nm).restype = C.c_double
            getattr(R, nm).argtypes = [C.c_double, C.c_double, C.c_double]
        R.ref_dipolar_prefactor.restype = C.c_double
        R.ref_dipolar_prefactor.argtypes = [C.c_double, C.c_double]
        R.ref_trapezoidal.restype = C.c_double
        R.ref_trapezoidal.argtypes = [_dp, _dp, C.c_size_t]
        R.ref_rng_mtnorm.argtypes = [C.c_ulong, C.c_double, _dp, C.c_size_t]
        R.ref_llg_drift.argtypes = [_dp, _dp, C.c_double, _dp]
        R.ref_llg_diffusion.argtypes = [_dp, _dp, C.c_double, C.c_double]
        R.ref_llg_drift_jacobian.argtypes = [_dp, _dp, C.c_double, _dp, _dp]
        R.ref_llg_diffusion_jacobian.argtypes = [_dp, _dp, C.c_double, C.c_double]
        R.ref_llg_multi_diffusion.argtypes = [_dp, _dp, _dp, _dp, C.c_size_t]
        R.ref_llg_multi_drift_quasijacobian.argtypes = [_dp, _dp, _dp, _dp, _dp, C.c_size_t]
        R.ref_llg_multi_diffusion_jacobian.argtypes = [_dp, _dp, _dp, _dp, C.c_size_t]
        R.ref_field_uniaxial.argtypes = [_dp, _dp, _dp]
        R.ref_field_multi_add_uniaxial.argtypes = [_dp, _dp, _dp, _dp, C.c_size_t]
        R.ref_field_multi_add_uniaxial_jacobian.argtypes = [_dp, _dp, _dp, C.c_size_t]
        R.ref_field_multi_add_dipolar.argtypes = [_dp, C.c_double, C.c_double, _dp, _dp, _dp,
                                                  _dp, C.c_size_t]
        R.ref_distances.argtypes = [_dp, C.c_size_t, _dp, _dp]
        R.ref_field_square_fourier.restype = C.c_double
        R.ref_field_square_fourier.argtypes = [C.c_double, C.c_double, C.c_double, C.c_size_t]
        R.ref_write_array.restype = C.c_int
        R.ref_write_array.argtypes = [C.c_char_p, _dp, C.c_size_t]
        _ref = R
    return _ref


def ptr(a):
    """Pointer to a contiguous numpy array (float64 / int32 / int64 / uint32)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    t = {np.dtype(np.float64): C.c_double, np.dtype(np.int32): C.c_int,
         np.dtype(np.int64): C.c_int64, np.dtype(np.uint32): C.c_uint32}[a.dtype]
    return a.ctypes.data_as(C.POINTER(t))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- RNG
def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32).copy()
    k = np.asarray(key, dtype=np.uint32).copy()
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(ptr(c), ptr(k), ptr(out))
    return out


def wiener_normals(seed, particle, j0, n, tag=0):
    out = np.zeros(n)
    lib().orc_wiener_normals(seed, particle, tag, j0, n, ptr(out))
    return out


def step_wiener(seed, n_particles, step, tag=0):
    """The 3N normals of one step (particle-major x,y,z), Philox stream."""
    w = np.zeros((n_particles, 3))
    for p in range(n_particles):
        w[p] = wiener_normals(seed, p, 3 * step, 3, tag)
    return w.reshape(-1)


# ---------------------------------------------------------------- schedule
def dnrm2(x):
    x = f64(x)
    return lib().orc_dnrm2(x.size, ptr(x))


def sample_schedule(dt_reduced, end_reduced, max_samples):
    k = np.zeros(max_samples, dtype=np.int64)
    n_steps = lib().orc_sample_schedule(dt_reduced, end_reduced, max_samples, ptr(k))
    return k, n_steps


# ---------------------------------------------------------------- full dynamics
def full_dynamics(radius, anisotropy, axes, m0, location, magnetisation, damping, temperature,
                  renorm, interactions, use_implicit, eps, time_step, end_time, max_samples,
                  seed, field_shape="constant", field_amplitude=0.0, field_frequency=0.0,
                  rng="mt", noise=None):
    """simulation::full_dynamics, physical overload (lib/simulation.cpp:476-624).

    Returns dict(time[S], field[S], x[N,S], y[N,S], z[N,S], stats)."""
    radius = f64(radius)
    n = radius.shape[0]
    anisotropy = f64(anisotropy)
    axes = f64(np.reshape(axes, (n, 3)))
    m0 = f64(np.reshape(m0, (n, 3)))
    location = f64(np.reshape(location, (n, 3)))
    S = int(max_samples)
    time = np.zeros(S)
    field = np.zeros(S)
    mx = np.zeros((n, S))
    my = np.zeros((n, S))
    mz = np.zeros((n, S))
    kind = {"mt": RNG_MT, "array": RNG_ARRAY, "philox": RNG_PHILOX}[rng]
    nz = f64(noise) if noise is not None else None
    st = Stats()
    rc = lib().orc_full_dynamics(
        ptr(radius), ptr(anisotropy), ptr(axes), ptr(m0), ptr(location), n, magnetisation,
        damping, temperature, int(renorm), int(interactions), int(use_implicit), eps,
        time_step, end_time, S, int(seed), FIELD[field_shape], field_amplitude,
        field_frequency, kind, ptr(nz), 0 if nz is None else nz.size, ptr(time), ptr(field),
        ptr(mx), ptr(my), ptr(mz), C.byref(st))
    if rc != 0:
        raise ValueError("orc_full_dynamics failed with code %d" % rc)
    return dict(time=time, field=field, x=mx, y=my, z=mz, stats=st)


def reduced_parameters(radius, anisotropy, magnetisation, damping, temperature, time_step,
                       end_time, field_amplitude=0.0, field_frequency=0.0):
    """The SI -> reduced conversion of lib/simulation.cpp:501-549 in numpy,
    written in the same evaluation order as the C restatement."""
    radius = np.asarray(radius, dtype=np.float64)
    n = radius.size
    vol = 4.0 / 3.0 * np.pi * radius * radius * radius
    vsum = 0.0
    for v in vol:
        vsum += v
    vav = vsum / n
    ksum = 0.0
    for k in np.asarray(anisotropy, dtype=np.float64):
        ksum += k
    kav = ksum / n
    hk = 2 * kav / 1.25663706e-6 / magnetisation
    tf = 1.76086e11 * 1.25663706e-6 * hk / (1 + damping * damping)
    sr = np.sqrt(damping * 1.38064852e-23 * temperature / (kav * vol) / (1 + damping * damping))
    return dict(volume=vol, v_av=vav, k_av=kav, hk=hk, time_factor=tf, dt=time_step * tf,
                end=end_time * tf, sr=sr, happ=field_amplitude / hk,
                fapp=field_frequency / tf)


# ---------------------------------------------------------------- pathwise sensitivities
def path_problem(axis, m0, dt, sig, alpha, happ, fdt, seeds, n_steps, sample_steps, tag=0):
    """orc_path_problem over P particles; dt/sig/alpha/happ are scalars or [P].
    Returns (struct, keep-alive arrays)."""
    axis = f64(np.reshape(axis, (-1, 3)))
    P = axis.shape[0]
    per = [f64(np.broadcast_to(np.asarray(x, dtype=np.float64), (P,))) for x in
           (dt, sig, alpha, happ)]
    m0 = f64(np.reshape(m0, (P, 3)))
    sd = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
    ks = np.ascontiguousarray(np.asarray(sample_steps, dtype=np.int64))
    pr = PathProblem(P, ptr(axis), ptr(m0), ptr(per[0]), ptr(per[1]), ptr(per[2]), ptr(per[3]),
                     float(fdt), sd.ctypes.data_as(C.POINTER(C.c_uint64)), tag, int(n_steps),
                     ptr(ks) if ks.size else None, ks.size)
    return pr, (axis, m0, per, sd, ks)


def path_run(pr):
    """(mz_sum [S], mz_final_sum, final_state [P,3]) of orc_path_run."""
    S, P = pr[0].n_samples, pr[0].n_particles
    mz = np.zeros(S)
    fin = C.c_double()
    st = np.zeros((P, 3))
    lib().orc_path_run(C.byref(pr[0]), ptr(mz) if S else None, C.byref(fin), ptr(st))
    return mz, fin.value, st


def path_tangent(pr, dq):
    """dq [P, D, 4] (derivatives of dt, sig, alpha, happ along D directions)
    -> (dmz_sum [S, D], dmz_final_sum [D])."""
    dq = f64(dq)
    P, D = dq.shape[0], dq.shape[1]
    S = pr[0].n_samples
    out = np.zeros((S, D))
    fin = np.zeros(D)
    lib().orc_path_tangent(C.byref(pr[0]), D, ptr(dq), ptr(out) if S else None, ptr(fin))
    return out, fin


# ---------------------------------------------------------------- discrete orientation model
def rk45(ode, x, t, h, tol):
    """One adaptive step of orc_rk45 with a Python ODE f(x, t) -> dx.
    Returns (next_state, k1..k6, t_new, h_new)."""
    n = len(x)
    x = f64(x)

    def cb(_ctx, dx, xx, tt):
        v = ode(np.ctypeslib.as_array(xx, shape=(n,)).copy(), tt)
        for i in range(n):
            dx[i] = v[i]

    fn = ODE_FN(cb)
    nxt, tmp = np.zeros(n), np.zeros(n)
    ks = [np.zeros(n) for _ in range(6)]
    hp, tp = C.c_double(h), C.c_double(t)
    lib().orc_rk45(ptr(nxt), ptr(tmp), *[ptr(k) for k in ks], C.byref(hp), C.byref(tp),
                   ptr(x), fn, None, n, tol)
    return nxt, ks, tp.value, hp.value


def dom_ensemble_dynamics(volume, anisotropy, temperature, magnetisation, alpha, field_shape,
                          h_reduced, frequency, initial_probs, time_step, end_time, max_samples,
                          n_components=1):
    """orc_dom_ensemble_dynamics -> dict(time, field (reduced), mz)."""
    S = int(max_samples)
    time, field, mz = np.zeros(S), np.zeros(S), np.zeros(S)
    shape = {"sine": 0, "square": 1, "constant": 2, "square_f": FIELD_SQUARE_F}[field_shape]
    lib().orc_dom_ensemble_dynamics(volume, anisotropy, temperature, magnetisation, alpha, shape,
                                    h_reduced, frequency, n_components, ptr(f64(initial_probs)),
                                    time_step, end_time, S, ptr(time), ptr(field), ptr(mz))
    return dict(time=time, field=field, mz=mz)
