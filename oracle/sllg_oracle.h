/*
 * sllg_oracle.h -- CPU FP64 restatement of owlas/magpy's stochastic LLG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path (magpy_amd/).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Functions that
 * belong to this build (the Philox Wiener stream) cite their published
 * algorithm instead.
 */
#ifndef SLLG_ORACLE_H
#define SLLG_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/constants.hpp:10-12 */
extern const double ORC_KB, ORC_MU0, ORC_GYROMAG;

/* ---------------- random numbers ---------------- */

/* Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy as
 * 1, 2, 3").  Counter-based: out = bijection_key(ctr). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Map a 32-bit word to the open interval (0,1): (x + 1/2) * 2^-32. */
double orc_u01(uint32_t x);

/* Wiener stream of one particle.  Normal number j of the stream of particle
 * `particle` of a member whose seed is `seed` is component (j & 3) of the
 * Box-Muller transform of Philox4x32-10 with
 *   key = (seed & 0xffffffff, seed >> 32)
 *   ctr = (c & 0xffffffff, c >> 32, particle, tag),  c = j >> 2.
 * Box-Muller pairs (w0,w1) -> (r cos 2 pi u(w1), r sin 2 pi u(w1)),
 * r = sqrt(-2 ln u(w0)), and likewise (w2,w3).  Writes n normals starting
 * at index j0. */
void orc_wiener_normals(uint64_t seed, uint32_t particle, uint32_t tag,
                        uint64_t j0, size_t n, double *out);

/* std::mt19937_64 + std::normal_distribution<double> (libstdc++), the
 * generator of RngMtNorm (lib/rng.cpp:14-24, include/rng.hpp:36-37). */
typedef struct {
    uint64_t mt[312];
    int mti;
    int has_saved;
    double saved;
    double mean, stddev;
} orc_mtnorm;
void orc_mtnorm_init(orc_mtnorm *g, uint64_t seed, double stddev);
uint64_t orc_mt64_next(orc_mtnorm *g);
double orc_mtnorm_get(orc_mtnorm *g);

/* ---------------- LLG terms (lib/llg.cpp) ---------------- */
void orc_llg_drift(double *deriv, const double *m, double alpha, const double *h);
void orc_llg_diffusion(double *deriv, const double *m, double sr, double alpha);
void orc_llg_drift_jacobian(double *jac, const double *m, double a,
                            const double *h, const double *hj);
void orc_llg_diffusion_jacobian(double *jac, const double *m, double sr, double alpha);
void orc_llg_multi_drift(double *deriv, const double *m, const double *alphas,
                         const double *h, size_t n);
void orc_llg_multi_diffusion(double *deriv, const double *m, const double *srs,
                             const double *alphas, size_t n);
void orc_llg_multi_drift_quasijacobian(double *jac, const double *m, const double *alphas,
                                       const double *h, const double *hj, size_t n);
void orc_llg_multi_diffusion_jacobian(double *jac, const double *m, const double *srs,
                                      const double *alphas, size_t n);

/* ---------------- fields (lib/field.cpp) ---------------- */
double orc_field_sinusoidal(double t, double h, double f);
double orc_field_square(double t, double h, double f);
void orc_field_uniaxial(double *h, const double *m, const double *axis);
void orc_field_uniaxial_jacobian(double *jac, const double *axis);
void orc_field_multi_add_uniaxial(double *h, const double *m, const double *axes,
                                  const double *k_red, size_t n);
void orc_field_multi_add_uniaxial_jacobian(double *jac, const double *axes,
                                           const double *k_red, size_t n);
void orc_field_multi_add_applied_z(double *h, double happ, size_t n);
double orc_dipolar_prefactor(double ms, double k_av);
void orc_dipolar_add_p2p_term(double *out, double vj, double rij3, const double *mj,
                              const double *dist, double prefactor);
void orc_field_multi_add_dipolar(double *h, double ms, double k_av, const double *v_red,
                                 const double *m, const double *dists,
                                 const double *dist_cubes, size_t n);

/* ---------------- distances (lib/distances.cpp) ---------------- */
void orc_pair_distance_vectors(double *out, const double *points, size_t n);
void orc_pair_distance_magnitudes(double *out, const double *vecs, size_t n);
void orc_pair_unit_vectors(double *out, const double *vecs, const double *mags, size_t n);

/* ---------------- integrators ---------------- */
/* sde(ctx, drift[n], diffusion[n*w], state[n], t) */
typedef void (*orc_sde_fn)(void *ctx, double *drift, double *diffusion,
                           const double *state, double t);
/* sde_jac(ctx, a[n], b[n*w], a'[n*n], b'[n*w*n], state[n], t_a, t_b) */
typedef void (*orc_sdejac_fn)(void *ctx, double *a, double *b, double *adash,
                              double *bdash, const double *state, double ta, double tb);
/* f(ctx, fout[n], jac[n*n], x[n]) */
typedef void (*orc_fj_fn)(void *ctx, double *fout, double *jac, const double *x);

void orc_heun(double *next, double *drift, double *trial_drift, double *diff,
              double *trial_diff, const double *cur, const double *w,
              orc_sde_fn sde, void *ctx, size_t n, size_t wdim, double t, double dt);
int orc_dgesv(int n, double *a, int *ipiv, double *b);
int orc_newton_noinv(double *x_root, double *x_tmp, double *jac, int *ipiv, int *lapack_err,
                     orc_fj_fn fj, void *ctx, const double *x0, int dim, double eps,
                     size_t max_iter);
int orc_implicit_midpoint(double *x, double *dwm, double *a_work, double *b_work,
                          double *adash_work, double *bdash_work, double *x_guess,
                          double *x_opt_tmp, double *x_opt_jac, int *ipiv,
                          const double *x0, const double *dw, orc_sdejac_fn sde, void *ctx,
                          size_t n, size_t wdim, double t, double dt, double eps,
                          size_t max_iter, int *newton_iters);

double orc_trapezoidal(const double *x, const double *y, size_t n);

/* ---------------- full dynamics (lib/simulation.cpp) ---------------- */
enum { ORC_FIELD_SINE = 0, ORC_FIELD_SQUARE = 1, ORC_FIELD_CONSTANT = 2 };
enum { ORC_RNG_MT = 0, ORC_RNG_ARRAY = 1, ORC_RNG_PHILOX = 2 };

typedef struct {
    int kind;
    /* ORC_RNG_MT */
    orc_mtnorm mt;
    /* ORC_RNG_ARRAY: RngArray (lib/rng.cpp:67-94) */
    const double *arr;
    size_t arr_len, arr_stride, arr_i;
    int exhausted;
    /* ORC_RNG_PHILOX: 3N draws per step, particle-major */
    uint64_t seed;
    uint32_t tag, n_particles;
    uint64_t step;
    uint32_t idx;
} orc_rng;

void orc_rng_init_mt(orc_rng *r, uint64_t seed, double stddev);
void orc_rng_init_array(orc_rng *r, const double *arr, size_t len, size_t stride);
void orc_rng_init_philox(orc_rng *r, uint64_t seed, uint32_t tag, uint32_t n_particles);
double orc_rng_get(orc_rng *r);

/* Per-run diagnostics the reference discards (lib/simulation.cpp:368-369). */
typedef struct {
    int64_t n_steps;
    int64_t newton_iters;      /* total over the run */
    int64_t newton_failures;   /* steps whose Newton returned non-zero */
    int rng_exhausted;
} orc_stats;

/* Reduced-unit overload, lib/simulation.cpp:132-429.  Arrays:
 * sr, k_red, v_red [n]; axes, m0 [3n]; unit_dists [n*n*3]; red_dists [n*n];
 * outputs time, field [S]; mx, my, mz [n*S] (particle-major).  Returns 0 on
 * success, -1 on bad arguments. */
int orc_full_dynamics_reduced(
    const double *sr, const double *k_red, const double *v_red, const double *axes,
    const double *m0, const double *unit_dists, const double *red_dists, size_t n,
    int field_shape, double happ, double fapp, double average_anisotropy,
    double average_volume, double damping, double ms, double dt, double end_time,
    orc_rng *rng, int renorm, int interactions, int use_implicit, double eps,
    int max_samples, double *time, double *field, double *mx, double *my, double *mz,
    orc_stats *stats);

/* Physical overload, lib/simulation.cpp:476-624 (seeded RNG replaced by the
 * caller's choice: rng_kind ORC_RNG_MT uses `seed` as the reference does,
 * ORC_RNG_PHILOX uses the build's Philox stream with the same seed,
 * ORC_RNG_ARRAY replays `noise`). */
int orc_full_dynamics(
    const double *radius, const double *anisotropy, const double *axes, const double *m0,
    const double *location, size_t n, double magnetisation, double damping,
    double temperature, int renorm, int interactions, int use_implicit, double eps,
    double time_step, double end_time, size_t max_samples, int64_t seed,
    int field_shape, double field_amplitude, double field_frequency,
    int rng_kind, const double *noise, size_t noise_len,
    double *time, double *field, double *mx, double *my, double *mz, orc_stats *stats);

/* BLAS dnrm2: Euclidean norm of x[0..n), overflow-free, within one ulp. */
double orc_dnrm2(int n, const double *x);

/* Sample schedule of the reduced loop (lib/simulation.cpp:342-406): k[s] is
 * the step index whose state is recorded as sample s; returns the number of
 * integration steps the loop takes (k[S-1] + 1). */
int64_t orc_sample_schedule(double dt, double end_time, size_t max_samples, int64_t *k);


/* ---------------- pathwise sensitivities ----------------
 * The differentiable Heun path of independent particles (one particle per
 * member, k_red = v_red = 1).  Step n of particle p:
 *   v1 = dt h(m, s_{n+1}) + sig w_n,     m~ = m + L(m, v1)
 *   v2 = dt h(m~, s_{n+2}) + sig w_n,    m' = m + (L(m, v1) + L(m~, v2)) / 2
 *   h(m, s) = (m.e) e + happ s z,  L(m, v) = -m x v + alpha (v |m|^2 - m (m.v))
 *   s_k = sin(2 pi frac(k fdt)),  fdt = f * dt in SI (field cycles per step)
 *   w_n = normals 3n, 3n+1, 3n+2 of the Philox stream (seed_p, particle 0, tag).
 * This is integrator::heun (lib/integrators.cpp:372-405) called at
 * t = (n+1) dt (lib/simulation.cpp:355, 372-375) with the drift and
 * diffusion of lib/llg.cpp:14-29, 92-106 (a dt + B w sqrt(dt) = L(m, h dt +
 * sr sqrt(dt) w)) and the fields of lib/field.cpp:37-40, 116-132; the phase
 * of the sine is taken in SI cycles so that it does not move with the
 * parameters.  The objective reads the ensemble sums of m_z at the steps
 * sample_steps[s] (state after that many steps) and after n_steps. */
typedef struct {
    size_t n_particles;
    const double *axis;          /* [P][3] easy axes */
    const double *m0;            /* [P][3] initial state */
    const double *dt;            /* [P] reduced step */
    const double *sig;           /* [P] sr sqrt(dt) */
    const double *alpha;         /* [P] damping */
    const double *happ;          /* [P] reduced field amplitude */
    double fdt;                  /* field cycles per step */
    const uint64_t *seed;        /* [P] */
    uint32_t tag;
    int64_t n_steps;
    const int64_t *sample_steps; /* [S], ascending, each in [0, n_steps] */
    size_t n_samples;
} orc_path_problem;

/* sin(2 pi frac(k fdt)) */
double orc_path_phase(double fdt, int64_t k);

/* One step of the path above (m updated in place). */
void orc_path_step(double m[3], const double e[3], const double w[3], double dt, double sig,
                   double alpha, double happ, double s1, double s2);

/* Forward run: mz_sum[s] = sum_p m_z,p(sample_steps[s]), *mz_final_sum =
 * sum_p m_z,p(n_steps); final_state [P][3] may be NULL. */
void orc_path_run(const orc_path_problem *pr, double *mz_sum, double *mz_final_sum,
                  double *final_state);

/* Forward-mode tangent of the run along n_dirs directions: dq[(p*n_dirs +
 * d)*4 + j] is the derivative of particle p's (dt, sig, alpha, happ)[j]
 * along direction d.  Writes dmz_sum[s*n_dirs + d] = d/dtheta_d of
 * mz_sum[s] and dmz_final_sum[d]. */
void orc_path_tangent(const orc_path_problem *pr, size_t n_dirs, const double *dq,
                      double *dmz_sum, double *dmz_final_sum);


/* ---------------- discrete orientation model (two-state master equation) ---------------- */
/* field::square_fourier (lib/field.cpp:67-78) */
double orc_field_square_fourier(double t, double h, double f, size_t n_components);
/* dom::transition_matrix (lib/dom.cpp:30-56): Neel-Brown rates, W row-major 2x2 */
void orc_dom_transition_matrix(double *W, double k, double v, double T, double h, double ms,
                               double alpha);
/* stochastic::master_equation (lib/stochastic_processes.cpp:18-25): derivs = W x */
void orc_master_equation(double *derivs, const double *W, const double *x, size_t dim);
/* integrator::rk45 (lib/integrators.cpp:152-251), Cash-Karp tableau
 * (include/integrators.hpp:128-160), with the ODE as a callback */
typedef void (*orc_ode_fn)(void *ctx, double *dx, const double *x, double t);
void orc_rk45(double *next_state, double *temp_state, double *k1, double *k2, double *k3,
              double *k4, double *k5, double *k6, double *h_ptr, double *t_ptr,
              const double *current_state, orc_ode_fn derivs, void *ctx, size_t n_dims,
              double tol);
/* simulation::dom_ensemble_dynamics (lib/simulation.cpp:659-767) with the
 * reduced field h(t) of core.simulate_dom (magpy/core.pyx:205-249):
 * field_shape ORC_FIELD_* or ORC_FIELD_SQUARE_F (3), amplitude already
 * reduced by H_k.  Writes time, reduced field and mz = p0 - p1 at the
 * max_samples regular samples (first-order hold). */
#define ORC_FIELD_SQUARE_F 3
void orc_dom_ensemble_dynamics(double volume, double anisotropy, double temperature,
                               double magnetisation, double alpha, int field_shape,
                               double h_reduced, double frequency, size_t n_components,
                               const double initial_probs[2], double time_step, double end_time,
                               size_t max_samples, double *time, double *field, double *mz);

#ifdef __cplusplus
}
#endif
#endif
