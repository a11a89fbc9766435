# cython: language_level=3, boundscheck=False, wraparound=False
"""Cython binding of libmagpy_amd.so (include/magpy_amd.h).

Mirrors owlas/magpy's magpy/core.pyx: ``simulate`` keeps the signature and
result dictionary of magpy/core.pyx:128-203 and ``get_KB/get_mu0/get_gamma``
those of :29-34.  ``simulate_ensemble`` is the batched call behind
``EnsembleModel.simulate``.  The GIL is released around the device run.
"""
import numpy as np

from libc.stdint cimport int32_t, int64_t, uint32_t, uint64_t


cdef extern from "magpy_amd.h" nogil:
    ctypedef struct mp_ensemble_spec:
        int64_t n_members
        const int64_t *member_offsets
        const double *radius
        const double *anisotropy
        const double *anisotropy_axis
        const double *magnetisation_direction
        const double *location
        const double *magnetisation
        const double *damping
        const double *temperature
        const int32_t *field_shape
        const double *field_amplitude
        const double *field_frequency
        const int64_t *seed
        double time_step
        double end_time
        int64_t max_samples
        int32_t renorm
        int32_t interactions
        int32_t scheme
        int32_t precision
        double implicit_tol
        int32_t implicit_max_iter
        int32_t noise
        const double *noise_array
        int64_t noise_steps
        uint32_t noise_tag
        int32_t device
        int32_t dipolar
        int64_t start_step
        const double *start_state

    ctypedef struct mp_ensemble_out:
        double *time
        double *field
        double *member_time
        double *member_field
        double *mx
        double *my
        double *mz
        double *ens_m
        double *final_state
        uint32_t *status
        int64_t n_steps
        int64_t newton_iters
        int64_t newton_failures
        double kernel_ms

    int mp_ensemble_run(const mp_ensemble_spec *spec, mp_ensemble_out *out)
    ctypedef struct mp_session:
        pass
    int mp_session_create(const mp_ensemble_spec *spec, int want_trajectories, int want_ensemble,
                          mp_session **session)
    int mp_session_advance(mp_session *session, int64_t n_steps, int64_t *steps_done)
    int mp_session_info(mp_session *session, int64_t *n_steps_total, int64_t *launches,
                        double *kernel_ms, int32_t *cluster_block, int32_t *dipolar,
                        int32_t *cluster_lanes)
    int mp_dipolar_field(const double *location, const double *radius, const double *anisotropy,
                         size_t n_particles, double magnetisation, const double *m, size_t reps,
                         int mode, double *h, int device)
    int mp_session_state(mp_session *session, double *state, int64_t *step)
    int mp_session_finish(mp_session *session, mp_ensemble_out *out)
    void mp_session_destroy(mp_session *session)
    int mp_sample_schedule(double dt_reduced, double end_reduced, int64_t max_samples,
                           int64_t *k, int64_t *n_steps)
    int mp_wiener_normals(uint64_t seed, uint32_t particle, uint32_t tag, uint64_t j0, size_t n,
                          int precision, double *out, int device)
    ctypedef struct mp_adjoint_spec:
        int64_t n_particles
        const double *radius_shape
        const double *anisotropy_axis
        const double *magnetisation_direction
        const int64_t *seed
        double magnetisation
        double temperature
        double field_frequency
        double time_step
        int64_t n_steps
        const int64_t *sample_steps
        int64_t n_samples
        int64_t checkpoint_interval
        int32_t precision
        uint32_t noise_tag
        int32_t device
    ctypedef struct mp_adjoint:
        pass
    int mp_adjoint_create(const mp_adjoint_spec *spec, mp_adjoint **out)
    void mp_adjoint_destroy(mp_adjoint *a)
    int mp_adjoint_begin(mp_adjoint *a, const double *theta, int record_gradient)
    int mp_adjoint_forward_advance(mp_adjoint *a, int64_t max_steps, int64_t *steps_done)
    int mp_adjoint_forward_result(mp_adjoint *a, double *ens_mz, double *final_mz,
                                  uint32_t *status)
    int mp_adjoint_reverse_begin(mp_adjoint *a, const double *grad_ens_mz, double grad_final_mz)
    int mp_adjoint_reverse_advance(mp_adjoint *a, int64_t max_steps, int64_t *steps_left)
    int mp_adjoint_reverse_result(mp_adjoint *a, double *grad_theta)
    int mp_adjoint_particle_gradients(mp_adjoint *a, double *grad)
    int mp_adjoint_info(mp_adjoint *a, double *forward_ms, double *reverse_ms,
                        int64_t *checkpoint_bytes, int64_t *scratch_bytes)
    int mp_adjoint_reduced(const mp_adjoint_spec *spec, const double *theta, double *q,
                           double *dq)
    ctypedef struct mp_dom_spec:
        int64_t n_particles
        const double *volume
        const double *anisotropy
        const double *temperature
        const double *magnetisation
        const double *damping
        const double *initial_probabilities
        int32_t field_shape
        double field_amplitude
        double field_frequency
        int64_t field_n_components
        double time_step
        double end_time
        int64_t max_samples
        int64_t max_steps
        int32_t device
    int mp_dom_run(const mp_dom_spec *spec, double *time, double *field, double *mz,
                   uint32_t *status, int64_t *steps, double *kernel_ms)
    int mp_save_results(const char *prefix, const double *mx, const double *my,
                        const double *mz, const double *field, const double *time,
                        size_t n_samples)
    void mp_constants(double *kb, double *mu0, double *gamma)
    const char *mp_last_error()
    const char *mp_version()
    int mp_device_info(int device, char *buf, size_t length)

    int MP_OK, MP_ERR_ARG, MP_ERR_HIP, MP_ERR_NODEV, MP_ERR_UNSUPPORTED, MP_ERR_ALLOC
    int MP_FIELD_SINE, MP_FIELD_SQUARE, MP_FIELD_CONSTANT
    int MP_SCHEME_HEUN, MP_SCHEME_IMPLICIT_MIDPOINT
    int MP_PRECISION_FP32, MP_PRECISION_FP64
    int MP_NOISE_PHILOX, MP_NOISE_ARRAY
    int MP_DIPOLAR_AUTO, MP_DIPOLAR_PAIRS, MP_DIPOLAR_GEMM_BF16X3, MP_DIPOLAR_GEMM_F32


class MagpyAMDError(RuntimeError):
    """A non-zero status from libmagpy_amd.so."""

    def __init__(self, code, message):
        super().__init__("magpy_amd error %d: %s" % (code, message))
        self.code = code


class UnsupportedError(MagpyAMDError):
    pass


FIELD_SHAPES = {'sine': MP_FIELD_SINE, 'square': MP_FIELD_SQUARE, 'constant': MP_FIELD_CONSTANT}
PRECISIONS = {'fp32': MP_PRECISION_FP32, 'fp64': MP_PRECISION_FP64}
DIPOLAR = {'auto': MP_DIPOLAR_AUTO, 'pairs': MP_DIPOLAR_PAIRS,
           'gemm_bf16x3': MP_DIPOLAR_GEMM_BF16X3, 'gemm_f32': MP_DIPOLAR_GEMM_F32}
DIPOLAR_NAMES = {v: k for k, v in DIPOLAR.items()}


def _check(int rc):
    if rc != MP_OK:
        msg = mp_last_error().decode()
        if rc == MP_ERR_UNSUPPORTED:
            raise UnsupportedError(rc, msg)
        if rc == MP_ERR_ARG:
            raise ValueError(msg)
        raise MagpyAMDError(rc, msg)


cpdef get_KB():
    cdef double kb
    mp_constants(&kb, NULL, NULL)
    return kb


cpdef get_mu0():
    cdef double mu0
    mp_constants(NULL, &mu0, NULL)
    return mu0


cpdef get_gamma():
    cdef double g
    mp_constants(NULL, NULL, &g)
    return g


def version():
    return mp_version().decode()


def device_info(int device=0):
    cdef char buf[256]
    _check(mp_device_info(device, buf, 256))
    return buf.decode()


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def sample_schedule(double dt_reduced, double end_reduced, int64_t max_samples):
    """Step index recorded as each sample, and the steps taken (lib/simulation.cpp:342-406)."""
    k = np.zeros(max_samples, dtype=np.int64)
    cdef int64_t[::1] kv = k
    cdef int64_t n_steps = 0
    _check(mp_sample_schedule(dt_reduced, end_reduced, max_samples, &kv[0], &n_steps))
    return k, n_steps


def dipolar_field(location, radius, anisotropy, double magnetisation, m, mode='gemm_bf16x3',
                  int device=0):
    """Reduced dipolar field of states m [reps, N, 3] of one cluster, from the
    device product of the dipole tensor with the states (mode 'gemm_bf16x3',
    'gemm_f32', or 'pairs' for the FP64 product).  Returns [reps, N, 3]."""
    loc = _c(np.reshape(location, (-1, 3)), np.float64)
    cdef size_t n = loc.shape[0]
    r = _c(radius, np.float64)
    k = _c(anisotropy, np.float64)
    mm = _c(np.reshape(m, (-1, n, 3)), np.float64)
    cdef size_t reps = mm.shape[0]
    h = np.zeros((reps, n, 3))
    cdef const double[:, ::1] loc_v = loc
    cdef const double[::1] r_v = r, k_v = k
    cdef const double[:, :, ::1] m_v = mm
    cdef double[:, :, ::1] h_v = h
    cdef int md = DIPOLAR[mode]
    cdef int rc
    with nogil:
        rc = mp_dipolar_field(&loc_v[0, 0], &r_v[0], &k_v[0], n, magnetisation, &m_v[0, 0, 0],
                              reps, md, &h_v[0, 0, 0], device)
    _check(rc)
    return h


def wiener_normals(uint64_t seed, uint32_t particle, uint64_t j0, size_t n, uint32_t tag=0,
                   precision='fp64', int device=0):
    """Normals j0..j0+n-1 of one particle's Philox Wiener stream, generated on the device."""
    out = np.zeros(max(n, 1))
    cdef double[::1] ov = out
    cdef int prec = PRECISIONS[precision]
    cdef int rc
    with nogil:
        rc = mp_wiener_normals(seed, particle, tag, j0, n, prec, &ov[0], device)
    _check(rc)
    return out[:n]


cdef class _Spec:
    """An mp_ensemble_spec over numpy arrays it keeps alive."""
    cdef mp_ensemble_spec sp
    cdef public object arrays
    cdef public int64_t M, P, S

    def __init__(self, member_offsets, radius, anisotropy, anisotropy_axis,
                 magnetisation_direction, location, magnetisation, damping, temperature,
                 field_shape, field_amplitude, field_frequency, seeds, double time_step,
                 double end_time, int64_t max_samples, bint renorm=False, bint interactions=True,
                 bint use_implicit=False, double implicit_tol=1e-9, int implicit_max_iter=1000,
                 precision='fp64', noise=None, uint32_t noise_tag=0, int device=0,
                 dipolar='auto', int64_t start_step=0, start_state=None):
        if max_samples < 1:
            raise ValueError("max_samples must be >= 1 (-1 never terminates in the reference loop)")
        offs = _c(member_offsets, np.int64)
        M = offs.shape[0] - 1
        P = int(offs[M])
        r = _c(radius, np.float64)
        k = _c(anisotropy, np.float64)
        ax = _c(np.reshape(anisotropy_axis, (P, 3)), np.float64)
        m0 = _c(np.reshape(magnetisation_direction, (P, 3)), np.float64)
        loc = _c(np.reshape(location, (P, 3)), np.float64)
        ms = _c(np.broadcast_to(magnetisation, (M,)), np.float64)
        al = _c(np.broadcast_to(damping, (M,)), np.float64)
        tk = _c(np.broadcast_to(temperature, (M,)), np.float64)
        if isinstance(field_shape, str):
            field_shape = [field_shape] * M
        fs = _c([FIELD_SHAPES[v] if isinstance(v, str) else v for v in field_shape], np.int32)
        fa = _c(np.broadcast_to(field_amplitude, (M,)), np.float64)
        ff = _c(np.broadcast_to(field_frequency, (M,)), np.float64)
        sd = _c(np.broadcast_to(seeds, (M,)), np.int64)
        if r.shape[0] != P or k.shape[0] != P or fs.shape[0] != M:
            raise ValueError("array lengths do not match member_offsets")
        nz = None
        if noise is not None:
            nz = _c(noise, np.float64)
            if nz.ndim != 3 or nz.shape[1] != P or nz.shape[2] != 3:
                raise ValueError("noise must have shape [steps, P, 3]")
        st0 = None
        if start_state is not None:
            st0 = _c(np.reshape(start_state, (P, 3)), np.float64)
        self.arrays = (offs, r, k, ax, m0, loc, ms, al, tk, fs, fa, ff, sd, nz, st0)
        self.M, self.P, self.S = M, P, max_samples

        cdef const int64_t[::1] offs_v = offs
        cdef const double[::1] r_v = r, k_v = k, ms_v = ms, al_v = al, tk_v = tk
        cdef const double[::1] fa_v = fa, ff_v = ff
        cdef const double[:, ::1] ax_v = ax, m0_v = m0, loc_v = loc
        cdef const int32_t[::1] fs_v = fs
        cdef const int64_t[::1] sd_v = sd
        cdef const double[::1] nz_v
        self.sp.n_members = M
        self.sp.member_offsets = &offs_v[0]
        self.sp.radius = &r_v[0]
        self.sp.anisotropy = &k_v[0]
        self.sp.anisotropy_axis = &ax_v[0, 0]
        self.sp.magnetisation_direction = &m0_v[0, 0]
        self.sp.location = &loc_v[0, 0]
        self.sp.magnetisation = &ms_v[0]
        self.sp.damping = &al_v[0]
        self.sp.temperature = &tk_v[0]
        self.sp.field_shape = &fs_v[0]
        self.sp.field_amplitude = &fa_v[0]
        self.sp.field_frequency = &ff_v[0]
        self.sp.seed = &sd_v[0]
        self.sp.time_step = time_step
        self.sp.end_time = end_time
        self.sp.max_samples = max_samples
        self.sp.renorm = 1 if renorm else 0
        self.sp.interactions = 1 if interactions else 0
        self.sp.scheme = MP_SCHEME_IMPLICIT_MIDPOINT if use_implicit else MP_SCHEME_HEUN
        self.sp.precision = PRECISIONS[precision]
        self.sp.implicit_tol = implicit_tol
        self.sp.implicit_max_iter = implicit_max_iter
        self.sp.noise_tag = noise_tag
        self.sp.device = device
        self.sp.dipolar = DIPOLAR[dipolar]
        self.sp.start_step = start_step
        cdef const double[:, ::1] st_v
        if st0 is not None:
            st_v = st0
            self.sp.start_state = &st_v[0, 0]
        else:
            self.sp.start_state = NULL
        if nz is not None:
            nz_v = nz.reshape(-1)
            self.sp.noise = MP_NOISE_ARRAY
            self.sp.noise_array = &nz_v[0]
            self.sp.noise_steps = nz.shape[0]
        else:
            self.sp.noise = MP_NOISE_PHILOX
            self.sp.noise_array = NULL
            self.sp.noise_steps = 0


cdef object _outputs(mp_ensemble_out *out, int64_t M, int64_t P, int64_t S, trajectories,
                     ensemble_sum, member_times):
    """Allocate result arrays and point `out` at them."""
    time = np.zeros(S)
    field = np.zeros(S)
    fin = np.zeros((P, 3))
    status = np.zeros(P, dtype=np.uint32)
    cdef double[::1] time_v = time, field_v = field
    cdef double[:, ::1] fin_v = fin
    cdef uint32_t[::1] st_v = status
    cdef double[:, ::1] x_v, y_v, z_v, mt_v, mf_v, e_v
    out.time = &time_v[0]
    out.field = &field_v[0]
    out.member_time = NULL
    out.member_field = NULL
    out.mx = NULL
    out.my = NULL
    out.mz = NULL
    out.ens_m = NULL
    out.final_state = &fin_v[0, 0]
    out.status = &st_v[0]
    res = {'N': P, 'time': time, 'field': field, 'final_state': fin, 'status': status}
    if trajectories:
        x = np.zeros((P, S))
        y = np.zeros((P, S))
        z = np.zeros((P, S))
        x_v = x
        y_v = y
        z_v = z
        out.mx = &x_v[0, 0]
        out.my = &y_v[0, 0]
        out.mz = &z_v[0, 0]
        res['x'], res['y'], res['z'] = x, y, z
    if member_times:
        mt = np.zeros((M, S))
        mf = np.zeros((M, S))
        mt_v = mt
        mf_v = mf
        out.member_time = &mt_v[0, 0]
        out.member_field = &mf_v[0, 0]
        res['member_time'], res['member_field'] = mt, mf
    if ensemble_sum:
        ens = np.zeros((S, 3))
        e_v = ens
        out.ens_m = &e_v[0, 0]
        res['ensemble_m'] = ens
    return res


cdef _stats(mp_ensemble_out *out, res):
    res['n_steps'] = out.n_steps
    res['newton_iters'] = out.newton_iters
    res['newton_failures'] = out.newton_failures
    res['kernel_ms'] = out.kernel_ms


def simulate_ensemble(member_offsets, radius, anisotropy, anisotropy_axis,
                      magnetisation_direction, location, magnetisation, damping, temperature,
                      field_shape, field_amplitude, field_frequency, seeds,
                      double time_step, double end_time, int64_t max_samples,
                      bint renorm=False, bint interactions=True, bint use_implicit=False,
                      double implicit_tol=1e-9, int implicit_max_iter=1000,
                      precision='fp64', trajectories=True, ensemble_sum=True,
                      noise=None, uint32_t noise_tag=0, int device=0, member_times=False,
                      dipolar='auto', int64_t start_step=0, start_state=None):
    """Run every member of an ensemble in one batched device call.

    Per-particle arrays are concatenated over members in member order;
    ``member_offsets`` [M+1] is the CSR of particles per member.  Per-member
    arrays have length M.  Returns a dict with 'time', 'field' [S] (member
    0), optional 'member_time'/'member_field' [M,S], 'x','y','z' [P,S]
    (Ms*m, A/m) when ``trajectories``, 'ensemble_m' [S,3] (sum over members
    and particles of Ms*m) when ``ensemble_sum``, 'final_state' [P,3],
    'status' [P], 'n_steps', 'newton_iters', 'newton_failures', 'kernel_ms'.
    ``noise`` (optional) is an array [steps, P, 3] of Wiener draws replacing
    the Philox stream.  ``start_step``/``start_state`` [P,3] resume a run at
    that step from a 'final_state' or Session.state() of an earlier run of
    the same ensemble (the Philox stream is indexed by step, so a split run
    equals the unsplit one); samples before start_step come back as NaN.
    """
    cdef _Spec spec = _Spec(member_offsets, radius, anisotropy, anisotropy_axis,
                            magnetisation_direction, location, magnetisation, damping,
                            temperature, field_shape, field_amplitude, field_frequency, seeds,
                            time_step, end_time, max_samples, renorm, interactions,
                            use_implicit, implicit_tol, implicit_max_iter, precision, noise,
                            noise_tag, device, dipolar, start_step, start_state)
    cdef mp_ensemble_out out
    res = _outputs(&out, spec.M, spec.P, spec.S, trajectories, ensemble_sum, member_times)
    cdef int rc
    with nogil:
        rc = mp_ensemble_run(&spec.sp, &out)
    _check(rc)
    _stats(&out, res)
    return res


cdef class Session:
    """A resident ensemble run (mp_session_*): advance it in blocks of steps,
    read its state, then finish() for the same dict simulate_ensemble returns.
    Arguments as simulate_ensemble."""
    cdef mp_session *s
    cdef _Spec spec
    cdef object traj, ens

    def __cinit__(self):
        self.s = NULL

    def __init__(self, *args, trajectories=False, ensemble_sum=True, **kw):
        self.spec = _Spec(*args, **kw)
        self.traj = bool(trajectories)
        self.ens = bool(ensemble_sum)
        cdef int rc
        cdef mp_session *ptr = NULL
        cdef int wt = 1 if self.traj else 0
        cdef int we = 1 if self.ens else 0
        with nogil:
            rc = mp_session_create(&self.spec.sp, wt, we, &ptr)
        _check(rc)
        self.s = ptr

    def __dealloc__(self):
        if self.s != NULL:
            mp_session_destroy(self.s)
            self.s = NULL

    def advance(self, int64_t n_steps):
        """Advance by up to n_steps (synchronous); returns steps done so far."""
        cdef int64_t done = 0
        cdef int rc
        with nogil:
            rc = mp_session_advance(self.s, n_steps, &done)
        _check(rc)
        return done

    def info(self):
        cdef int64_t tot = 0, launches = 0
        cdef double ms = 0
        cdef int32_t cblock = 0, dip = 0, clanes = 0
        _check(mp_session_info(self.s, &tot, &launches, &ms, &cblock, &dip, &clanes))
        return {'n_steps': tot, 'launches': launches, 'kernel_ms': ms, 'cluster_block': cblock,
                'cluster_lanes': clanes,
                'dipolar': DIPOLAR_NAMES.get(dip, 'none') if dip else 'none'}

    def state(self):
        st = np.zeros((self.spec.P, 3))
        cdef double[:, ::1] v = st
        cdef int64_t step = 0
        _check(mp_session_state(self.s, &v[0, 0], &step))
        return st, step

    def finish(self, member_times=False):
        cdef mp_ensemble_out out
        res = _outputs(&out, self.spec.M, self.spec.P, self.spec.S, self.traj, self.ens,
                       member_times)
        cdef int rc
        with nogil:
            rc = mp_session_finish(self.s, &out)
        _check(rc)
        _stats(&out, res)
        return res


cpdef simulate(radius, anisotropy, anisotropy_axis, magnetisation_direction, location,
               double magnetisation, double damping, double temperature, bint renorm,
               bint interactions, bint use_implicit, double time_step, double end_time,
               int max_samples, long seed, str field_shape='constant',
               double field_amplitude=0.0, double field_frequency=0.0,
               double implicit_tol=1e-9, precision='fp64', int device=0):
    """simulation::full_dynamics for one cluster, as magpy/core.pyx:128-203.

    Returns {'N', 'time', 'field', 'x': {i: arr}, 'y': {...}, 'z': {...}}.
    """
    radius = _c(radius, np.float64)
    n = radius.shape[0]
    res = simulate_ensemble(
        [0, n], radius, anisotropy, anisotropy_axis, magnetisation_direction, location,
        magnetisation, damping, temperature, [field_shape], field_amplitude, field_frequency,
        [seed], time_step, end_time, max_samples, renorm=renorm, interactions=interactions,
        use_implicit=use_implicit, implicit_tol=implicit_tol, precision=precision,
        trajectories=True, ensemble_sum=False, device=device)
    return {
        'N': n,
        'time': res['time'],
        'field': res['field'],
        'x': {i: res['x'][i] for i in range(n)},
        'y': {i: res['y'][i] for i in range(n)},
        'z': {i: res['z'][i] for i in range(n)},
    }


# ------------------------------------------------------------------ sensitivities
cdef class _AdjointSpec:
    """mp_adjoint_spec with the arrays it points to."""
    cdef mp_adjoint_spec sp
    cdef object arrays

    def __init__(self, radius_shape, anisotropy_axis, magnetisation_direction, seeds,
                 double magnetisation, double temperature, double field_frequency,
                 double time_step, int64_t n_steps, sample_steps, int64_t checkpoint_interval=1000,
                 precision='fp32', uint32_t noise_tag=0, int device=0):
        shp = _c(radius_shape, np.float64).reshape(-1)
        P = shp.shape[0]
        if P == 0:
            raise ValueError("no particles")
        ax = _c(np.reshape(anisotropy_axis, (P, 3)), np.float64)
        m0 = _c(np.reshape(magnetisation_direction, (P, 3)), np.float64)
        sd = _c(seeds, np.int64).reshape(-1)
        ks = _c(sample_steps, np.int64).reshape(-1)
        if sd.shape[0] != P:
            raise ValueError("one seed per particle")
        self.arrays = (shp, ax, m0, sd, ks)
        cdef const double[::1] shp_v = shp
        cdef const double[:, ::1] ax_v = ax, m0_v = m0
        cdef const int64_t[::1] sd_v = sd
        cdef const int64_t[::1] ks_v
        self.sp.n_particles = P
        self.sp.radius_shape = &shp_v[0]
        self.sp.anisotropy_axis = &ax_v[0, 0]
        self.sp.magnetisation_direction = &m0_v[0, 0]
        self.sp.seed = &sd_v[0]
        self.sp.magnetisation = magnetisation
        self.sp.temperature = temperature
        self.sp.field_frequency = field_frequency
        self.sp.time_step = time_step
        self.sp.n_steps = n_steps
        self.sp.n_samples = ks.shape[0]
        if ks.shape[0]:
            ks_v = ks
            self.sp.sample_steps = &ks_v[0]
        else:
            self.sp.sample_steps = NULL
        self.sp.checkpoint_interval = checkpoint_interval
        self.sp.precision = PRECISIONS[precision]
        self.sp.noise_tag = noise_tag
        self.sp.device = device


def adjoint_reduced(theta, radius_shape, double magnetisation, double temperature,
                    double time_step):
    """Reduced parameters q [P, 4] = (dt, sr sqrt(dt), alpha, happ) of every
    particle at theta = (K, alpha, H, median radius) and their derivatives
    dq [P, 4, 4] (dq[p, j, i] = dq_j / dtheta_i); host-side, no device."""
    shp = _c(radius_shape, np.float64).reshape(-1)
    P = shp.shape[0]
    cdef _AdjointSpec spec = _AdjointSpec(shp, np.zeros((P, 3)), np.zeros((P, 3)),
                                          np.zeros(P, np.int64), magnetisation, temperature,
                                          0.0, time_step, 1, [])
    th = _c(theta, np.float64)
    q = np.zeros((P, 4))
    dq = np.zeros((P, 4, 4))
    cdef const double[::1] th_v = th
    cdef double[:, ::1] q_v = q
    cdef double[:, :, ::1] dq_v = dq
    _check(mp_adjoint_reduced(&spec.sp, &th_v[0], &q_v[0, 0], &dq_v[0, 0, 0]))
    return q, dq


cdef class Adjoint:
    """Forward and reverse passes of the pathwise sensitivities
    (mp_adjoint_*).  theta = (anisotropy K, damping alpha, field amplitude H,
    median radius).  forward(theta) -> (ens_mz [S], final_mz, status);
    reverse(grad_ens_mz, grad_final_mz) -> dJ/dtheta [4] for the last
    recording forward pass.  Both also exist as begin/advance/result steps."""
    cdef mp_adjoint *a
    cdef _AdjointSpec spec
    cdef readonly int64_t n_particles, n_steps, n_samples
    cdef readonly int64_t forward_passes

    def __cinit__(self):
        self.a = NULL

    def __init__(self, *args, **kw):
        self.spec = _AdjointSpec(*args, **kw)
        cdef mp_adjoint *ptr = NULL
        cdef int rc
        with nogil:
            rc = mp_adjoint_create(&self.spec.sp, &ptr)
        _check(rc)
        self.a = ptr
        self.n_particles = self.spec.sp.n_particles
        self.n_steps = self.spec.sp.n_steps
        self.n_samples = self.spec.sp.n_samples
        self.forward_passes = 0

    def __dealloc__(self):
        if self.a != NULL:
            mp_adjoint_destroy(self.a)
            self.a = NULL

    def begin(self, theta, record_gradient=True):
        """Start a forward pass at theta; record_gradient=False computes
        values only (no reverse pass possible)."""
        th = _c(theta, np.float64)
        if th.shape[0] != 4:
            raise ValueError("theta has four entries (K, alpha, H, radius)")
        cdef const double[::1] th_v = th
        _check(mp_adjoint_begin(self.a, &th_v[0], 1 if record_gradient else 0))
        self.forward_passes += 1

    def forward_advance(self, int64_t n):
        cdef int64_t done = 0
        cdef int rc
        with nogil:
            rc = mp_adjoint_forward_advance(self.a, n, &done)
        _check(rc)
        return done

    def forward_result(self):
        ens = np.zeros(max(self.n_samples, 1))
        cdef double[::1] ev = ens
        cdef double fin = 0
        cdef uint32_t st = 0
        cdef int rc
        with nogil:
            rc = mp_adjoint_forward_result(self.a, &ev[0], &fin, &st)
        _check(rc)
        return ens[:self.n_samples], fin, st

    def forward(self, theta, record_gradient=True):
        self.begin(theta, record_gradient)
        return self.forward_result()

    def reverse_begin(self, grad_ens_mz, double grad_final_mz):
        g = _c(grad_ens_mz, np.float64).reshape(-1)
        if g.shape[0] != self.n_samples:
            raise ValueError("grad_ens_mz needs one entry per sample")
        if g.shape[0] == 0:
            g = np.zeros(1)
        cdef const double[::1] gv = g
        _check(mp_adjoint_reverse_begin(self.a, &gv[0], grad_final_mz))

    def reverse_advance(self, int64_t n):
        cdef int64_t left = 0
        cdef int rc
        with nogil:
            rc = mp_adjoint_reverse_advance(self.a, n, &left)
        _check(rc)
        return left

    def reverse_result(self):
        grad = np.zeros(4)
        cdef double[::1] gv = grad
        cdef int rc
        with nogil:
            rc = mp_adjoint_reverse_result(self.a, &gv[0])
        _check(rc)
        return grad

    def reverse(self, grad_ens_mz, double grad_final_mz):
        self.reverse_begin(grad_ens_mz, grad_final_mz)
        return self.reverse_result()

    def particle_gradients(self):
        """[P, 4]: each particle's part of the last reverse pass's gradient."""
        out = np.zeros((self.n_particles, 4))
        cdef double[:, ::1] ov = out
        cdef int rc
        with nogil:
            rc = mp_adjoint_particle_gradients(self.a, &ov[0, 0])
        _check(rc)
        return out

    def info(self):
        cdef double fms = 0, rms = 0
        cdef int64_t ckb = 0, scb = 0
        _check(mp_adjoint_info(self.a, &fms, &rms, &ckb, &scb))
        return {'forward_ms': fms, 'reverse_ms': rms, 'checkpoint_bytes': ckb,
                'scratch_bytes': scb}


def save_results(str prefix, mx, my, mz, field, time):
    """simulation::save_results (lib/simulation.cpp:38-63): raw FP64 files
    prefix.mx/.my/.mz/.field/.time of one particle's samples."""
    arrs = [_c(a, np.float64).reshape(-1) for a in (mx, my, mz, field, time)]
    n = arrs[0].shape[0]
    if any(a.shape[0] != n for a in arrs):
        raise ValueError("all five arrays need the same length")
    if n == 0:
        arrs = [np.zeros(1) for _ in arrs]
    cdef const double[::1] a0 = arrs[0], a1 = arrs[1], a2 = arrs[2], a3 = arrs[3], a4 = arrs[4]
    b = prefix.encode()
    cdef const char *pb = b
    _check(mp_save_results(pb, &a0[0], &a1[0], &a2[0], &a3[0], &a4[0], n))


DOM_FIELD_SHAPES = {'sine': 0, 'square': 1, 'constant': 2, 'square_f': 3}


def simulate_dom_ensemble(volume, anisotropy, temperature, magnetisation, damping,
                          initial_probabilities, double time_step, double end_time,
                          int64_t max_samples, str field_shape='constant',
                          double field_amplitude=0.0, double field_frequency=0.0,
                          int64_t field_n_components=1, int64_t max_steps=0, int device=0):
    """Discrete orientation model of every particle of an ensemble in one
    device launch (mp_dom_run).  Per-particle arrays (length P); returns
    dict(time [S], field [S] (A/m), mz [P, S], status [P], steps [P],
    kernel_ms (device time of the integration kernel))."""
    v = _c(volume, np.float64).reshape(-1)
    P = v.shape[0]
    if P == 0:
        raise ValueError("no particles")
    per = [_c(np.broadcast_to(np.asarray(x, dtype=np.float64), (P,)), np.float64)
           for x in (anisotropy, temperature, magnetisation, damping)]
    ip = _c(np.broadcast_to(np.asarray(initial_probabilities, dtype=np.float64), (P, 2)),
            np.float64)
    if field_shape not in DOM_FIELD_SHAPES:
        raise ValueError("field_shape must be one of %s" % sorted(DOM_FIELD_SHAPES))
    cdef const double[::1] v_v = v, k_v = per[0], t_v = per[1], m_v = per[2], a_v = per[3]
    cdef const double[:, ::1] ip_v = ip
    cdef mp_dom_spec sp
    sp.n_particles = P
    sp.volume = &v_v[0]
    sp.anisotropy = &k_v[0]
    sp.temperature = &t_v[0]
    sp.magnetisation = &m_v[0]
    sp.damping = &a_v[0]
    sp.initial_probabilities = &ip_v[0, 0]
    sp.field_shape = DOM_FIELD_SHAPES[field_shape]
    sp.field_amplitude = field_amplitude
    sp.field_frequency = field_frequency
    sp.field_n_components = field_n_components
    sp.time_step = time_step
    sp.end_time = end_time
    sp.max_samples = max_samples
    sp.max_steps = max_steps
    sp.device = device
    S = max(max_samples, 1)
    time = np.zeros(S)
    field = np.zeros(S)
    mz = np.zeros((P, S))
    status = np.zeros(P, dtype=np.uint32)
    steps = np.zeros(P, dtype=np.int64)
    cdef double[::1] tv = time, fv = field
    cdef double[:, ::1] mzv = mz
    cdef uint32_t[::1] sv = status
    cdef int64_t[::1] stv = steps
    cdef int rc
    cdef double kms = 0.0
    with nogil:
        rc = mp_dom_run(&sp, &tv[0], &fv[0], &mzv[0, 0], &sv[0], &stv[0], &kms)
    _check(rc)
    return {'time': time, 'field': field, 'mz': mz, 'status': status, 'steps': steps,
            'kernel_ms': kms}


cpdef simulate_dom(initial_probabilities, double volume, double anisotropy, double temperature,
                   double magnetisation, double alpha, double time_step, double end_time,
                   size_t max_samples, str field_shape, double field_amplitude,
                   double field_frequency, size_t field_n_components):
    """core.simulate_dom (magpy/core.pyx:205-262): one particle's state
    probabilities; the result dictionary of the reference (mz in 'z', x and y
    zero)."""
    r = simulate_dom_ensemble([volume], [anisotropy], [temperature], [magnetisation], [alpha],
                              np.reshape(np.asarray(initial_probabilities, dtype=np.float64),
                                         (1, 2)),
                              time_step, end_time, max_samples, field_shape, field_amplitude,
                              field_frequency, field_n_components)
    if r['status'][0]:
        raise FloatingPointError("the master equation did not integrate (non-finite rates or "
                                 "no accepted step)")
    S = max_samples
    return {'N': 1, 'time': r['time'], 'field': r['field'], 'x': {0: np.zeros(S)},
            'y': {0: np.zeros(S)}, 'z': {0: r['mz'][0]}}
