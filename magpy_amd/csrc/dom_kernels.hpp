// dom_kernels.hpp -- discrete orientation model of independent particles.
//
// Each particle is a two-state system (up, down) whose probabilities obey
// the master equation dp/dt = W(h(t)) p with the Neel-Brown rates of
// dom::transition_matrix (lib/dom.cpp:30-56); the reference integrates it
// per particle with the adaptive Cash-Karp RK45 of integrator::rk45
// (lib/integrators.cpp:152-251) inside simulation::dom_ensemble_dynamics
// (lib/simulation.cpp:659-767), sampling mz = p_up - p_down by first-order
// hold.  Here one lane integrates one particle in FP64, all particles of an
// ensemble in one launch; samples are written [sample][particle].
//
// The reference loops until a step is accepted and until time passes each
// sample; a NaN error estimate or a vanishing step would spin forever.
// Each lane bounds both (kDomMaxRetries per step, max_steps per lane) and
// flags the particle instead, so every wave finishes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mp {

constexpr int kDomMaxRetries = 200;

struct DomArgs {
    int n;
    const double *k, *v, *temp, *ms, *alpha;   // per particle, SI
    const double *happ;                       // per particle: field amplitude / H_k
    const double *p_up, *p_down;              // initial probabilities
    int shape;                                // 0 sine, 1 square, 2 constant, 3 square Fourier
    double freq;
    int64_t n_components;
    double time_step, end_time;
    int64_t S;
    int64_t max_steps;
    double *mz;                               // [S][n]
    uint32_t *status;                         // bit 0: gave up (non-finite or no progress)
    int64_t *steps;                           // accepted steps per particle
};

__device__ inline double dom_field(const DomArgs &a, double h, double t)
{
#pragma clang fp contract(off)
    const double pi = 3.14159265358979323846;
    switch (a.shape) {
    case 0: return h * sin(2 * pi * a.freq * t);                       // lib/field.cpp:37-40
    case 1: return h * ((long long)(t * a.freq * 2) % 2 ? -1.0 : 1.0);  // lib/field.cpp:51-54
    case 3: {                                                          // lib/field.cpp:67-78
        double f = 0.0;
        for (int64_t k = 1; k < a.n_components + 1; ++k) {
            const double odd = (double)(2 * k - 1);
            f += sin(2 * pi * odd * a.freq * t) / odd;
        }
        return f * (4 / pi * h);
    }
    default: return h;
    }
}

// The particle's constants of dom::transition_matrix (lib/dom.cpp:30-56):
// sigma = K V / kT and the field-free part tau_N sqrt(pi) / sigma^1.5 of the
// prefactor, in the reference's order of operations.
struct DomParticle {
    double happ, sigma, c0;
};

__device__ inline DomParticle dom_particle(double k, double v, double temp, double ms, double al,
                                           double happ)
{
#pragma clang fp contract(off)
    const double kb = 1.38064852e-23, gamma = 1.76086e11, pi = 3.14159265358979323846;
    const double sigma = k * v / kb / temp;
    const double taun = v * ms * (1 + al * al) / 2.0 / gamma / al / kb / temp;
    return DomParticle{happ, sigma, taun * sqrt(pi) / pow(sigma, 1.5)};
}

// The rates at time t: W = [[-r2, r1], [r2, -r1]], pre = c0 / (1 - h^2)
// (the reference's taun sqrt(pi) / pow(sigma, 1.5) / (1 - h h), left to
// right).  They depend on t (through the field) and the particle, not on the
// probabilities.
__device__ inline void dom_rates(const DomArgs &a, const DomParticle &q, double t, double &r1,
                                 double &r2)
{
#pragma clang fp contract(off)
    const double h = dom_field(a, q.happ, t);
    const double e1 = q.sigma * (1 - h) * (1 - h);
    const double e2 = q.sigma * (1 + h) * (1 + h);
    const double pre = q.c0 / (1 - h * h);
    r1 = 1.0 / pre * (1 - h) * exp(-e1);
    r2 = 1.0 / pre * (1 + h) * exp(-e2);
}

// dp = W p
__device__ inline void dom_apply(double r1, double r2, const double *p, double *dp)
{
#pragma clang fp contract(off)
    dp[0] = -r2 * p[0] + r1 * p[1];
    dp[1] = r2 * p[0] + -r1 * p[1];
}

// integrator::rk45 (lib/integrators.cpp:152-251) with the Cash-Karp tableau
// of include/integrators.hpp:128-160; returns false if no step was accepted
// within kDomMaxRetries attempts.  The rates of an attempt's six stage times
// are formed before its stages (they do not depend on the stage values), so
// their FP64 sin/exp evaluations can overlap instead of alternating with
// the stage arithmetic; (r1t, r2t) are the rates at t, which do not change
// between attempts and on return hold the rates at the new t (the accepted
// attempt's fifth stage time, t + h * 1.0 = t + h).  The same values, in
// the same expressions, as evaluating each stage's rates in turn.
__device__ inline bool dom_rk45(const DomArgs &a, const DomParticle &q, double *next,
                                const double *x, double *h_ptr, double *t_ptr, double tol,
                                double &r1t, double &r2t)
{
#pragma clang fp contract(off)
    double k1[2], k2[2], k3[2], k4[2], k5[2], k6[2], y[2], x5[2];
    double h = *h_ptr, t = *t_ptr, err = 0.0;
    double r1[5], r2[5];     // stages 2..6
    bool ok = false;
    dom_apply(r1t, r2t, x, k1);
    for (int attempt = 0; attempt < kDomMaxRetries && !ok; ++attempt) {
        const double ts[5] = {t + h * 0.2, t + h * 0.3, t + h * 0.6, t + h * 1.0,
                              t + h * (7.0 / 8.0)};
#pragma unroll
        for (int st = 0; st < 5; ++st) dom_rates(a, q, ts[st], r1[st], r2[st]);
        for (int i = 0; i < 2; ++i) y[i] = k1[i] * h * 0.2 + x[i];
        dom_apply(r1[0], r2[0], y, k2);
        for (int i = 0; i < 2; ++i) y[i] = x[i] + h * (3.0 / 40.0 * k1[i] + 9.0 / 40.0 * k2[i]);
        dom_apply(r1[1], r2[1], y, k3);
        for (int i = 0; i < 2; ++i)
            y[i] = x[i] + h * (3.0 / 10.0 * k1[i] + -9.0 / 10.0 * k2[i] + 6.0 / 5.0 * k3[i]);
        dom_apply(r1[2], r2[2], y, k4);
        for (int i = 0; i < 2; ++i)
            y[i] = x[i] + h * (-11.0 / 54.0 * k1[i] + 2.5 * k2[i] + -70.0 / 27.0 * k3[i] +
                               35.0 / 27.0 * k4[i]);
        dom_apply(r1[3], r2[3], y, k5);
        for (int i = 0; i < 2; ++i)
            y[i] = x[i] + h * (1631.0 / 55296.0 * k1[i] + 175.0 / 512.0 * k2[i] +
                               575.0 / 13824.0 * k3[i] + 44275.0 / 110592.0 * k4[i] +
                               253.0 / 4096.0 * k5[i]);
        dom_apply(r1[4], r2[4], y, k6);
        for (int i = 0; i < 2; ++i)
            x5[i] = x[i] + h * (37.0 / 378.0 * k1[i] + 250.0 / 621.0 * k3[i] +
                                125.0 / 594.0 * k4[i] + 512.0 / 1771.0 * k6[i]);
        for (int i = 0; i < 2; ++i)
            next[i] = x[i] + h * (2825.0 / 27648.0 * k1[i] + 18575.0 / 48384.0 * k3[i] +
                                  13525.0 / 55296.0 * k4[i] + 277.0 / 14336.0 * k5[i] +
                                  0.25 * k6[i]);
        // the reference's pow(., 2) and pow(., 0.5) of these norms as a square
        // and a square root: x * x is pow(x, 2) exactly, and sqrt, correctly
        // rounded, is what an accurate pow(x, 0.5) returns; a general FP64
        // pow is a long dependent chain, and these were five of the six per
        // attempt
        const double mag = sqrt(x[0] * x[0] + x[1] * x[1]);
        const double d0 = fabs(x5[0] - next[0]), d1 = fabs(x5[1] - next[1]);
        err = sqrt(d0 * d0 + d1 * d1);
        err /= (2.0 * tol * (1 + mag));
        if (err < 1.0) {
            ok = true;
        } else {
            double f = 0.84 * pow(err, -0.2);
            f = fabs(f) < 0.1 ? 0.1 : f;
            h *= f;
        }
    }
    if (!ok) return false;
    *t_ptr = t + h;
    r1t = r1[3];
    r2t = r2[3];
    double f = err == 0.0 ? 5.0 : 0.84 * pow(err, -0.2);
    f = f > 5 ? 5.0 : f;
    *h_ptr = f * h;
    return true;
}

// simulation::dom_ensemble_dynamics (lib/simulation.cpp:659-767), one lane
// per particle.
__global__ __launch_bounds__(64) void dom_kernel(DomArgs a)
{
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const double k = a.k[p], v = a.v[p], temp = a.temp[p], ms = a.ms[p], al = a.alpha[p];
    const double happ = a.happ[p];
    double next[2] = {a.p_up[p], a.p_down[p]}, last[2];
    a.mz[p] = next[0] - next[1];
    const double ts = a.end_time / (double)(a.S - 1);
    const double max_dt = a.end_time / 1000.0;
    double t = 0.0, t_last = 0.0, dt = 0.01 * a.time_step;
    const DomParticle q = dom_particle(k, v, temp, ms, al, happ);
    double r1t, r2t;     // rates at t
    dom_rates(a, q, t, r1t, r2t);
    int64_t steps = 0;
    bool alive = true;
    for (int64_t s = 1; s < a.S; ++s) {
        while (alive && t <= (double)s * ts) {
            last[0] = next[0];
            last[1] = next[1];
            t_last = t;
            if (++steps > a.max_steps ||
                !dom_rk45(a, q, next, last, &dt, &t, a.time_step, r1t, r2t)) {
                alive = false;
                break;
            }
            dt = dt > max_dt ? max_dt : dt;
        }
        double out;
        if (alive) {
            const double mz_last = last[0] - last[1], mz_next = next[0] - next[1];
            const double beta = (mz_next - mz_last) / (t - t_last);
            out = mz_last + beta * ((double)s * ts - t_last);
        } else {
            out = __builtin_nan("");
        }
        a.mz[(size_t)s * a.n + p] = out;
    }
    a.status[p] = alive ? 0u : 1u;
    a.steps[p] = steps;
}

}  // namespace mp
