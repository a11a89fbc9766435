/*
 * magpy_amd.h -- C ABI of the MI355X sLLG integrator (libmagpy_amd.so).
 *
 * Plain pointers and sizes only: no C++ types, no exceptions, no torch types.
 * Every entry returns MP_OK (0) or an MP_ERR_* code; mp_last_error() gives the
 * message of the calling thread's last failure.  Inputs are read-only and
 * caller-owned; outputs are caller-allocated (NULL = not wanted).
 *
 * Each entry cites the owlas/magpy interface it replaces.
 */
#ifndef MAGPY_AMD_H
#define MAGPY_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MP_ABI_VERSION 1

/* status codes */
#define MP_OK 0
#define MP_ERR_ARG 1          /* invalid argument (message says which)        */
#define MP_ERR_HIP 2          /* HIP runtime failure                          */
#define MP_ERR_NODEV 3        /* no HIP device visible                        */
#define MP_ERR_UNSUPPORTED 4  /* valid request this build does not run        */
#define MP_ERR_ALLOC 5        /* device or host allocation failed             */

/* applied-field shapes: the values of field::options (include/field.hpp:95-97) */
#define MP_FIELD_SINE 0
#define MP_FIELD_SQUARE 1
#define MP_FIELD_CONSTANT 2

#define MP_SCHEME_HEUN 0              /* integrator::heun, lib/integrators.cpp:372-405 */
#define MP_SCHEME_IMPLICIT_MIDPOINT 1 /* integrator::implicit_midpoint, :576-651 */

#define MP_PRECISION_FP32 0
#define MP_PRECISION_FP64 1

/* Wiener increments.  PHILOX: normal j of particle p (index within its
 * member) is component j%4 of Box-Muller over Philox4x32-10 with
 * key = (seed lo32, seed hi32), ctr = (j/4 lo32, j/4 hi32, p, noise_tag);
 * step n uses normals 3n, 3n+1, 3n+2 (x, y, z).  ARRAY: the caller's draws,
 * the RngArray replacement (lib/rng.cpp:67-94), laid out [step][particle][3]
 * over the whole ensemble. */
#define MP_NOISE_PHILOX 0
#define MP_NOISE_ARRAY 1

/* Dipolar field of interacting clusters.  PAIRS: per-lane pair sums of each
 * particle's cluster (any ensemble).  GEMM_BF16X3 / GEMM_F32: when every
 * member is the same cluster (positions, radii, anisotropy, axes, material,
 * field; only seeds and initial states differ), all replicas' fields at one
 * stage are one product Mt D^T on MFMA, in BF16 with the three-term split and
 * FP32 accumulation, or in FP32.  AUTO: GEMM_BF16X3 for FP32 Heun replica
 * ensembles of >= 64 members with >= 32 particles, PAIRS otherwise. */
#define MP_DIPOLAR_AUTO 0
#define MP_DIPOLAR_PAIRS 1
#define MP_DIPOLAR_GEMM_BF16X3 2
#define MP_DIPOLAR_GEMM_F32 3

/* per-particle status bits */
#define MP_STATUS_NONFINITE 1u      /* state became NaN/Inf                     */
#define MP_STATUS_NEWTON_FAIL 2u    /* an implicit step's Newton did not converge */

typedef struct mp_ensemble_spec {
    int64_t n_members;              /* clusters in the ensemble                  */
    const int64_t *member_offsets;  /* [n_members+1] CSR offsets into particles  */
    /* per particle (SI units, as simulation::full_dynamics takes them) */
    const double *radius;                  /* [P] m                              */
    const double *anisotropy;              /* [P] J/m^3                          */
    const double *anisotropy_axis;         /* [P*3] unit vectors                 */
    const double *magnetisation_direction; /* [P*3] initial unit vectors         */
    const double *location;                /* [P*3] m                            */
    /* per member */
    const double *magnetisation;    /* [M] A/m   */
    const double *damping;          /* [M]       */
    const double *temperature;      /* [M] K     */
    const int32_t *field_shape;     /* [M] MP_FIELD_* */
    const double *field_amplitude;  /* [M] A/m   */
    const double *field_frequency;  /* [M] Hz    */
    const int64_t *seed;            /* [M]       */
    /* shared by all members */
    double time_step;               /* s */
    double end_time;                /* s */
    int64_t max_samples;            /* >= 1; -1 (never terminates in the reference) is rejected */
    int32_t renorm;
    int32_t interactions;
    int32_t scheme;                 /* MP_SCHEME_* */
    int32_t precision;              /* MP_PRECISION_* (implicit midpoint runs in FP64) */
    double implicit_tol;            /* Newton relative tolerance (reference default 1e-9) */
    int32_t implicit_max_iter;      /* 1000 in the reference (lib/simulation.cpp:204) */
    int32_t noise;                  /* MP_NOISE_* */
    const double *noise_array;      /* [noise_steps][P][3] when noise == MP_NOISE_ARRAY */
    int64_t noise_steps;
    uint32_t noise_tag;
    int32_t device;                 /* HIP device ordinal */
    int32_t dipolar;                /* MP_DIPOLAR_*: how interacting clusters sum the field */
    /* Resume: start at step start_step from the reduced state start_state
     * [P][3] (NULL: magnetisation_direction).  The Wiener stream is keyed by
     * the step index, so a run split at any step and resumed gives the same
     * result.  Samples taken before start_step are returned as NaN. */
    int64_t start_step;
    const double *start_state;
} mp_ensemble_spec;

typedef struct mp_ensemble_out {
    /* sample times (s) and applied field (A/m) of member 0: [S] */
    double *time;
    double *field;
    /* the same for every member: [M][S] (optional) */
    double *member_time;
    double *member_field;
    /* per-particle magnetisation Ms*m (A/m), particle-major [P][S] as
     * simulation::results (include/simulation.hpp:32-48) */
    double *mx, *my, *mz;
    /* ensemble sum over members and particles of Ms*m per sample: [S][3] */
    double *ens_m;
    /* reduced state after the last step: [P][3] */
    double *final_state;
    /* per-particle MP_STATUS_* bits: [P] */
    uint32_t *status;
    /* filled on return */
    int64_t n_steps;            /* integration steps taken per member        */
    int64_t newton_iters;       /* total Newton iterations (implicit)        */
    int64_t newton_failures;    /* steps whose Newton hit max_iter or a singular matrix */
    double kernel_ms;           /* device time spent in the step kernels     */
} mp_ensemble_out;

/* Batched entry point: for every member, the semantics of
 * simulation::full_dynamics (physical overload, include/simulation.hpp:74-94,
 * lib/simulation.cpp:476-624) with that member's arrays and seed; replaces the
 * joblib fan-out of EnsembleModel.simulate (magpy/model.py:202-207). */
int mp_ensemble_run(const mp_ensemble_spec *spec, mp_ensemble_out *out);

/* mp_ensemble_run without device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code that must
 * not call hipMalloc, serving loops, harnesses that forbid allocation).  Same
 * arguments, semantics and results as mp_ensemble_run (the same kernels), but
 * every device buffer of the run (parameters, state, samples, ensemble
 * partial sums, field tables) is placed in the library's fixed scratch array
 * of 64 MiB per GPU (a static device array, held by each GPU the library runs
 * on until the process ends); launches go to the null stream, no stream or
 * event is created and the cluster layout is not autotuned (kernel_ms is the
 * host's clock around the synchronised launches).  A run that needs more
 * than the scratch returns MP_ERR_ALLOC with the sizes in mp_last_error()
 * before any kernel runs.  Calls that overlap on one GPU (from several host
 * threads) share the array and race; each call writes every buffer it reads,
 * so nothing of an earlier call's contents is used. */
int mp_ensemble_run_noalloc(const mp_ensemble_spec *spec, mp_ensemble_out *out);

/* One cluster, argument for argument simulation::full_dynamics (physical
 * overload, include/simulation.hpp:74-94) as core.simulate binds it
 * (magpy/core.pyx:72-93); outputs particle-major [n][max_samples] like the
 * vector<results> it returns.  Wiener increments from the Philox stream keyed
 * by `seed` (MP_NOISE_PHILOX, noise_tag 0). */
int mp_full_dynamics(const double *radius, const double *anisotropy,
                     const double *anisotropy_axis, const double *magnetisation_direction,
                     const double *location, size_t n_particles, double magnetisation,
                     double damping, double temperature, int renorm, int interactions,
                     int use_implicit, double eps, double time_step, double end_time,
                     size_t max_samples, long seed, int field_shape, double field_amplitude,
                     double field_frequency, int precision, double *time, double *field,
                     double *mx, double *my, double *mz);

/* Resident run: the ensemble's state and parameters stay in device memory
 * between calls, so a caller can advance it in blocks of steps (timing,
 * progress, checkpoints) and collect the same outputs mp_ensemble_run gives.
 * No counterpart in the reference, whose full_dynamics runs a member to the
 * end in one call (lib/simulation.cpp:342-406).  A session is used by one
 * thread at a time. */
typedef struct mp_session mp_session;
/* want_* select which outputs are recorded on the device */
int mp_session_create(const mp_ensemble_spec *spec, int want_trajectories, int want_ensemble,
                      mp_session **session);
/* advance every member by up to n_steps (stops at the end of the run);
 * *steps_done (optional) = steps taken so far after the call */
int mp_session_advance(mp_session *session, int64_t n_steps, int64_t *steps_done);
/* total steps of the run, kernel launches / device milliseconds so far, the
 * block size (lanes per workgroup) of interacting clusters (0 without pair
 * sums), the MP_DIPOLAR_* path the run uses (0 without interacting clusters)
 * and the lanes per particle of the pair-sum kernel (1 or 2; 0 without) */
int mp_session_info(mp_session *session, int64_t *n_steps_total, int64_t *launches,
                    double *kernel_ms, int32_t *cluster_block, int32_t *dipolar,
                    int32_t *cluster_lanes);
/* reduced state [P][3] and the step index it belongs to */
int mp_session_state(mp_session *session, double *state, int64_t *step);
/* run to the end and fill `out` as mp_ensemble_run would */
int mp_session_finish(mp_session *session, mp_ensemble_out *out);
void mp_session_destroy(mp_session *session);

/* Sample schedule of the reduced step loop (lib/simulation.cpp:342-406):
 * k[s] = step index recorded as sample s; *n_steps = steps the loop takes. */
int mp_sample_schedule(double dt_reduced, double end_reduced, int64_t max_samples,
                       int64_t *k, int64_t *n_steps);

/* Reduced dipolar field h[r][3i+a] of `reps` states m[r][3j+b] of one
 * cluster (positions, radii, anisotropy, Ms as simulation::full_dynamics
 * takes them), computed on the device as the product of the dipole tensor
 * with the states: mode MP_DIPOLAR_GEMM_BF16X3, MP_DIPOLAR_GEMM_F32, or
 * MP_DIPOLAR_PAIRS (FP64 product on the vector ALU).  Exposes the replica
 * path's product for accuracy checks; the reference computes the same field
 * per member in field::multi_add_dipolar (lib/field.cpp:187-203). */
int mp_dipolar_field(const double *location, const double *radius, const double *anisotropy,
                     size_t n_particles, double magnetisation, const double *m, size_t reps,
                     int mode, double *h, int device);

/* Fill out[n] with normals j0..j0+n-1 of one particle's Wiener stream
 * (MP_NOISE_PHILOX definition above) computed on the device in the given
 * precision.  Exposes the in-kernel generator for parity checks; no
 * counterpart in the reference (RngMtNorm::get, lib/rng.cpp:24, is what the
 * stream replaces). */
int mp_wiener_normals(uint64_t seed, uint32_t particle, uint32_t tag, uint64_t j0, size_t n,
                      int precision, double *out, int device);

/* ------------------------------------------------------------ sensitivities
 * Pathwise gradients of an ensemble of independent particles (one particle
 * per member, as EnsembleModel builds from one Model, magpy/model.py:146-208)
 * with respect to theta = (anisotropy K [J/m^3], damping alpha, field
 * amplitude H [A/m], median radius r [m]).  Particle p has radius
 * r * radius_shape[p] and its own seed; Ms, T, the sine field's frequency
 * and the time step are fixed.  The step is integrator::heun
 * (lib/integrators.cpp:372-405) in the reduced units of
 * lib/simulation.cpp:501-549, with the field's phase sin(2 pi f dt k) taken
 * at step k in SI units; the Wiener increments are the MP_NOISE_PHILOX
 * stream of (seed[p], particle 0, noise_tag) and do not depend on theta.
 * Outputs of a forward pass: ens_mz[s] = Ms * mean_p m_z,p after
 * sample_steps[s] steps (the reference's ensemble_magnetisation('z'),
 * magpy/results.py:134-151, at those steps) and final_mz = mean_p m_z,p
 * after n_steps.  A reverse pass returns the gradient of
 *   J = sum_s grad_ens_mz[s] * ens_mz[s] + grad_final_mz * final_mz
 * with respect to theta through the trajectories (the vector-Jacobian
 * product).  While a recording forward pass runs, the adjoint's linear maps
 * between sample steps are computed from states kept every
 * checkpoint_interval steps (recomputed in between) and held per sampling
 * interval in device memory (21 doubles per particle per interval), so the
 * reverse pass for any weights is a short scan.  No counterpart in the reference,
 * which has no derivatives; magpy_amd.adjoint wraps these calls in a
 * torch.autograd.Function. */
typedef struct mp_adjoint_spec {
    int64_t n_particles;
    const double *radius_shape;             /* [P] radius / median radius */
    const double *anisotropy_axis;          /* [P][3] */
    const double *magnetisation_direction;  /* [P][3] initial state */
    const int64_t *seed;                    /* [P] */
    double magnetisation;                   /* Ms [A/m] */
    double temperature;                     /* [K] */
    double field_frequency;                 /* [Hz] */
    double time_step;                       /* [s] */
    int64_t n_steps;
    const int64_t *sample_steps;            /* [S] ascending, in [0, n_steps] */
    int64_t n_samples;
    int64_t checkpoint_interval;            /* steps; a positive multiple of 4 */
    int32_t precision;                      /* MP_PRECISION_*: state and adjoint */
    uint32_t noise_tag;
    int32_t device;
} mp_adjoint_spec;

typedef struct mp_adjoint mp_adjoint;
int mp_adjoint_create(const mp_adjoint_spec *spec, mp_adjoint **out);
void mp_adjoint_destroy(mp_adjoint *a);
/* start a forward pass at theta (drops any earlier pass); record_gradient
 * = 0 runs the forward pass only (values, e.g. for finite differences) */
int mp_adjoint_begin(mp_adjoint *a, const double theta[4], int record_gradient);
/* advance the forward pass by up to max_steps; *steps_done = steps so far.
 * Returns when the forward kernels are done; the adjoint maps of the last
 * block may still run and overlap the next call's forward kernels */
int mp_adjoint_forward_advance(mp_adjoint *a, int64_t max_steps, int64_t *steps_done);
/* finish the forward pass and return its outputs (ens_mz [S], final_mz);
 * *status (optional) = OR over particles of MP_STATUS_NONFINITE */
int mp_adjoint_forward_result(mp_adjoint *a, double *ens_mz, double *final_mz, uint32_t *status);
/* start a reverse pass of the finished forward pass for the given weights */
int mp_adjoint_reverse_begin(mp_adjoint *a, const double *grad_ens_mz, double grad_final_mz);
/* run the reverse pass (max_steps is ignored: the scan is one launch);
 * *steps_left = 0 */
int mp_adjoint_reverse_advance(mp_adjoint *a, int64_t max_steps, int64_t *steps_left);
/* finish the reverse pass: grad_theta[4] = dJ/d(K, alpha, H, r) */
int mp_adjoint_reverse_result(mp_adjoint *a, double grad_theta[4]);
/* the same gradient split by particle: grad [P][4], the part of dJ/dtheta
 * that flows through each particle's trajectory (they sum to grad_theta) */
int mp_adjoint_particle_gradients(mp_adjoint *a, double *grad);
/* device milliseconds of the forward kernels and of the gradient kernels
 * (transfers, folding, scan) so far (waits for queued work), and the device
 * bytes held for checkpoints and segment scratch */
int mp_adjoint_info(mp_adjoint *a, double *forward_ms, double *reverse_ms, int64_t *checkpoint_bytes,
                    int64_t *scratch_bytes);
/* Reduced per-particle parameters q[p] = (dt, sr sqrt(dt), alpha, happ) at
 * theta and their derivatives dq[p][j][i] = dq_j / dtheta_i (host only) */
int mp_adjoint_reduced(const mp_adjoint_spec *spec, const double theta[4], double *q, double *dq);

/* ------------------------------------------------ discrete orientation model
 * simulation::dom_ensemble_dynamics (lib/simulation.cpp:659-767) as
 * core.simulate_dom calls it (magpy/core.pyx:205-262), for every particle of
 * an ensemble in one device launch: a two-state master equation with
 * Neel-Brown rates (lib/dom.cpp:30-93) integrated by adaptive Cash-Karp
 * RK45 (lib/integrators.cpp:152-251, tolerance = time_step), sampled at
 * max_samples regular times by first-order hold.  The field is applied along
 * each particle's axis; its amplitude (A/m) is reduced by each particle's
 * H_k = 2K / (mu0 Ms).  MP_FIELD_SQUARE_FOURIER is core.simulate_dom's
 * 'square_f' (field::square_fourier, lib/field.cpp:67-78). */
#define MP_FIELD_SQUARE_FOURIER 3
typedef struct mp_dom_spec {
    int64_t n_particles;
    const double *volume;                 /* [P] m^3 */
    const double *anisotropy;             /* [P] J/m^3 */
    const double *temperature;            /* [P] K */
    const double *magnetisation;          /* [P] A/m */
    const double *damping;                /* [P] */
    const double *initial_probabilities;  /* [P][2]: (up, down) */
    int32_t field_shape;                  /* MP_FIELD_* or MP_FIELD_SQUARE_FOURIER */
    double field_amplitude;               /* A/m */
    double field_frequency;               /* Hz */
    int64_t field_n_components;           /* square Fourier terms */
    double time_step;                     /* s: initial step / 100 and RK45 tolerance */
    double end_time;                      /* s */
    int64_t max_samples;                  /* >= 2 */
    int64_t max_steps;                    /* per particle; <= 0: 2^32.  A particle that
                                             reaches it (or a step that is never
                                             accepted) gets status 1 and NaN samples */
    int32_t device;
} mp_dom_spec;
/* time [S] (s), field [S] (A/m, particle 0's reduced field times its H_k, as
 * the reference returns it), mz [P][S] = p_up - p_down, status [P],
 * steps [P] accepted RK45 steps, kernel_ms the device time of the
 * integration kernel (HIP events around its launch); any output may be NULL */
int mp_dom_run(const mp_dom_spec *spec, double *time, double *field, double *mz,
               uint32_t *status, int64_t *steps, double *kernel_ms);

/* simulation::save_results (lib/simulation.cpp:38-63): writes prefix.mx,
 * prefix.my, prefix.mz, prefix.field and prefix.time, each the raw
 * native-endian FP64 array of n_samples values (io::write_array,
 * include/io.hpp:13-26), for one particle's results. */
int mp_save_results(const char *prefix, const double *mx, const double *my, const double *mz,
                    const double *field, const double *time, size_t n_samples);

/* get_KB / get_mu0 / get_gamma (magpy/core.pyx:29-34) */
void mp_constants(double *kb, double *mu0, double *gamma);

const char *mp_last_error(void);
const char *mp_version(void);
/* device name, CU count and clock of `device` into buf */
int mp_device_info(int device, char *buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
