/*
 * sllg_oracle.c -- CPU FP64 restatement of owlas/magpy's stochastic LLG path.
 *
 * TEST INFRASTRUCTURE ONLY (see sllg_oracle.h).  Reference citations are
 * relative to the owlas/magpy source tree.  Arithmetic is written in the
 * reference's evaluation order so that results agree bit for bit with the
 * reference's own building blocks compiled from source (oracle/_ref), and the
 * library is built with -ffp-contract=off for the same reason.
 *
 * Pinning status of each part is recorded in the test suite
 * (tests/test_oracle_*.py).  orc_dnrm2 is pinned against Python's math.hypot
 * (error below one ulp) on ordinary, huge and tiny vectors; the
 * full_dynamics orchestration (unit reduction and time factor, the field's
 * time convention, the zero-order-hold sample index, the SI conversion) is
 * pinned by closed-form solutions at zero temperature
 * (tests/test_oracle_closed_forms.py).
 */
#include "sllg_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

const double ORC_KB = 1.38064852e-23;   /* include/constants.hpp:10 */
const double ORC_MU0 = 1.25663706e-6;   /* include/constants.hpp:11 */
const double ORC_GYROMAG = 1.76086e11;  /* include/constants.hpp:12 */

/* ======================================================================
 * Philox4x32-10 and the Wiener stream
 * ==================================================================== */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double orc_u01(uint32_t x)
{
    return ((double)x + 0.5) * 0x1p-32;
}

static void orc_wiener_block(uint64_t seed, uint32_t particle, uint32_t tag, uint64_t c,
                             double z[4])
{
    uint32_t ctr[4] = {(uint32_t)c, (uint32_t)(c >> 32), particle, tag};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    orc_philox4x32_10(ctr, key, w);
    for (int p = 0; p < 2; ++p) {
        double r = sqrt(-2.0 * log(orc_u01(w[2 * p])));
        double th = 2.0 * M_PI * orc_u01(w[2 * p + 1]);
        z[2 * p] = r * cos(th);
        z[2 * p + 1] = r * sin(th);
    }
}

void orc_wiener_normals(uint64_t seed, uint32_t particle, uint32_t tag, uint64_t j0,
                        size_t n, double *out)
{
    double z[4];
    uint64_t have = UINT64_MAX;
    for (size_t i = 0; i < n; ++i) {
        uint64_t j = j0 + i;
        uint64_t c = j >> 2;
        if (c != have) {
            orc_wiener_block(seed, particle, tag, c, z);
            have = c;
        }
        out[i] = z[j & 3];
    }
}

/* ======================================================================
 * std::mt19937_64 (Matsumoto & Nishimura) and libstdc++'s
 * normal_distribution<double> (Marsaglia polar method over
 * generate_canonical<double,53>): the RngMtNorm generator,
 * lib/rng.cpp:14-24.
 * ==================================================================== */
#define MT_N 312
#define MT_M 156

void orc_mtnorm_init(orc_mtnorm *g, uint64_t seed, double stddev)
{
    g->mt[0] = seed;
    for (int i = 1; i < MT_N; ++i)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->mti = MT_N;
    g->has_saved = 0;
    g->saved = 0.0;
    g->mean = 0.0;
    g->stddev = stddev;
}

uint64_t orc_mt64_next(orc_mtnorm *g)
{
    const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
    const uint64_t mag = 0xB5026F5AA96619E9ULL;
    if (g->mti >= MT_N) {
        for (int i = 0; i < MT_N; ++i) {
            uint64_t y = (g->mt[i] & upper) | (g->mt[(i + 1) % MT_N] & lower);
            uint64_t v = g->mt[(i + MT_M) % MT_N] ^ (y >> 1);
            if (y & 1ULL) v ^= mag;
            g->mt[i] = v;
        }
        g->mti = 0;
    }
    uint64_t y = g->mt[g->mti++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= (y >> 43);
    return y;
}

/* generate_canonical<double, 53> over a 64-bit engine: one draw / 2^64. */
static double orc_mt_canonical(orc_mtnorm *g)
{
    double sum = (double)orc_mt64_next(g);
    double ret = sum / 18446744073709551616.0;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

double orc_mtnorm_get(orc_mtnorm *g)
{
    double ret;
    if (g->has_saved) {
        g->has_saved = 0;
        ret = g->saved;
    } else {
        double x, y, r2;
        do {
            x = 2.0 * orc_mt_canonical(g) - 1.0;
            y = 2.0 * orc_mt_canonical(g) - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        double mult = sqrt(-2 * log(r2) / r2);
        g->saved = x * mult;
        g->has_saved = 1;
        ret = y * mult;
    }
    return ret * g->stddev + g->mean;
}

/* ======================================================================
 * LLG drift and diffusion, lib/llg.cpp
 * ==================================================================== */

/* lib/llg.cpp:14-29 */
void orc_llg_drift(double *d, const double *m, double alpha, const double *h)
{
    d[0] = m[2] * h[1] - m[1] * h[2]
         + alpha * (h[0] * (m[1] * m[1] + m[2] * m[2]) - m[0] * (m[1] * h[1] + m[2] * h[2]));
    d[1] = m[0] * h[2] - m[2] * h[0]
         + alpha * (h[1] * (m[0] * m[0] + m[2] * m[2]) - m[1] * (m[0] * h[0] + m[2] * h[2]));
    d[2] = m[1] * h[0] - m[0] * h[1]
         + alpha * (h[2] * (m[0] * m[0] + m[1] * m[1]) - m[2] * (m[0] * h[0] + m[1] * h[1]));
}

/* lib/llg.cpp:92-106 (row-major 3x3, row = state component, col = Wiener) */
void orc_llg_diffusion(double *d, const double *m, double sr, double alpha)
{
    d[0] = alpha * sr * (m[1] * m[1] + m[2] * m[2]);
    d[1] = sr * (m[2] - alpha * m[0] * m[1]);
    d[2] = -sr * (m[1] + alpha * m[0] * m[2]);
    d[3] = -sr * (m[2] + alpha * m[0] * m[1]);
    d[4] = alpha * sr * (m[0] * m[0] + m[2] * m[2]);
    d[5] = sr * (m[0] - alpha * m[1] * m[2]);
    d[6] = sr * (m[1] - alpha * m[0] * m[2]);
    d[7] = -sr * (m[0] + alpha * m[1] * m[2]);
    d[8] = alpha * sr * (m[0] * m[0] + m[1] * m[1]);
}

/* lib/llg.cpp:68-81: d(drift)/dm given the field h and its Jacobian hj */
void orc_llg_drift_jacobian(double *j, const double *m, double a, const double *h,
                            const double *hj)
{
    j[0] = m[2] * hj[3] - m[1] * hj[6]
         + a * (-m[1] * h[1] - m[2] * h[2] + (m[1] * m[1] + m[2] * m[2]) * hj[0]
                - m[0] * (m[1] * hj[3] + m[2] * hj[6]));
    j[1] = -h[2] + m[2] * hj[4] - m[1] * hj[7]
         + a * (2 * m[1] * h[0] + (m[1] * m[1] + m[2] * m[2]) * hj[1]
                - m[0] * (h[1] + m[1] * hj[4] + m[2] * hj[7]));
    j[2] = h[1] + m[2] * hj[5] - m[1] * hj[8]
         + a * (2 * m[2] * h[0] + (m[1] * m[1] + m[2] * m[2]) * hj[2]
                - m[0] * (h[2] + m[1] * hj[5] + m[2] * hj[8]));
    j[3] = h[2] - m[2] * hj[0] + m[0] * hj[6]
         + a * (2 * m[0] * h[1] + (m[0] * m[0] + m[2] * m[2]) * hj[3]
                - m[1] * (h[0] + m[0] * hj[0] + m[2] * hj[6]));
    j[4] = -m[2] * hj[1] + m[0] * hj[7]
         + a * (-m[0] * h[0] - m[2] * h[2] + (m[0] * m[0] + m[2] * m[2]) * hj[4]
                - m[1] * (m[0] * hj[1] + m[2] * hj[7]));
    j[5] = -h[0] - m[2] * hj[2] + m[0] * hj[8]
         + a * (2 * m[2] * h[1] + (m[0] * m[0] + m[2] * m[2]) * hj[5]
                - m[1] * (h[2] + m[0] * hj[2] + m[2] * hj[8]));
    j[6] = -h[1] + m[1] * hj[0] - m[0] * hj[3]
         + a * (2 * m[0] * h[2] + (m[0] * m[0] + m[1] * m[1]) * hj[6]
                - m[2] * (h[0] + m[0] * hj[0] + m[1] * hj[3]));
    j[7] = h[0] + m[1] * hj[1] - m[0] * hj[4]
         + a * (2 * m[1] * h[2] + (m[0] * m[0] + m[1] * m[1]) * hj[7]
                - m[2] * (h[1] + m[0] * hj[1] + m[1] * hj[4]));
    j[8] = m[1] * hj[2] - m[0] * hj[5]
         + a * (-m[0] * h[0] - m[1] * h[1] + (m[0] * m[0] + m[1] * m[1]) * hj[8]
                - m[2] * (m[0] * hj[2] + m[1] * hj[5]));
}

/* lib/llg.cpp:118-158: jac[z + 3y + 9x] = d B[x][y] / d m[z] */
void orc_llg_diffusion_jacobian(double *jc, const double *m, double sr, double alpha)
{
    jc[0] = 0;
    jc[1] = 2 * alpha * sr * m[1];
    jc[2] = 2 * alpha * sr * m[2];
    jc[3] = -alpha * sr * m[1];
    jc[4] = -alpha * sr * m[2];
    jc[5] = sr;
    jc[6] = -alpha * sr * m[2];
    jc[7] = -sr;
    jc[8] = -alpha * sr * m[0];
    jc[9] = -alpha * sr * m[1];
    jc[10] = -alpha * sr * m[0];
    jc[11] = -sr;
    jc[12] = 2 * alpha * sr * m[0];
    jc[13] = 0;
    jc[14] = 2 * alpha * sr * m[2];
    jc[15] = sr;
    jc[16] = -alpha * sr * m[2];
    jc[17] = -alpha * sr * m[1];
    jc[18] = -alpha * sr * m[2];
    jc[19] = sr;
    jc[20] = -alpha * sr * m[0];
    jc[21] = -sr;
    jc[22] = -alpha * sr * m[2];
    jc[23] = -alpha * sr * m[1];
    jc[24] = 2 * alpha * sr * m[0];
    jc[25] = 2 * alpha * sr * m[2];
    jc[26] = 0;
}

/* lib/llg.cpp:257-266 */
void orc_llg_multi_drift(double *d, const double *m, const double *alphas, const double *h,
                         size_t n)
{
    for (size_t p = 0; p < n; ++p)
        orc_llg_drift(d + 3 * p, m + 3 * p, alphas[p], h + 3 * p);
}

/* lib/llg.cpp:296-324: dense (3N)x(3N) block-diagonal diffusion, zeros
 * written explicitly */
void orc_llg_multi_diffusion(double *d, const double *m, const double *srs,
                             const double *alphas, size_t n)
{
    const size_t n3 = 3 * n;
    double blk[9];
    for (size_t b = 0; b < n; ++b) {
        orc_llg_diffusion(blk, m + 3 * b, srs[b], alphas[b]);
        for (size_t r = 0; r < 3; ++r) {
            double *row = d + (3 * b + r) * n3;
            for (size_t c = 0; c < 3 * b; ++c) row[c] = 0.0;
            for (size_t c = 0; c < 3; ++c) row[3 * b + c] = blk[3 * r + c];
            for (size_t c = 3 * (b + 1); c < n3; ++c) row[c] = 0.0;
        }
    }
}

/* lib/llg.cpp:378-402.  The particle's field Jacobian is read as the nine
 * consecutive entries starting at hj + 3p of the (3N)x(3N) row-major field
 * Jacobian, exactly as the reference indexes it (for N > 1 this is not the
 * particle's diagonal block; it only changes Newton's convergence path). */
void orc_llg_multi_drift_quasijacobian(double *jac, const double *m, const double *alphas,
                                       const double *h, const double *hj, size_t n)
{
    double w[9];
    const size_t n3 = 3 * n;
    for (size_t p = 0; p < n; ++p) {
        orc_llg_drift_jacobian(w, m + 3 * p, alphas[p], h + 3 * p, hj + 3 * p);
        for (size_t r = 0; r < 3; ++r)
            for (size_t c = 0; c < 3; ++c)
                jac[3 * p + c + (3 * p + r) * n3] = w[c + r * 3];
    }
}

/* lib/llg.cpp:409-427: (3N)^3 tensor, only block-diagonal entries written */
void orc_llg_multi_diffusion_jacobian(double *jac, const double *m, const double *srs,
                                      const double *alphas, size_t n)
{
    double w[27];
    const size_t n3 = 3 * n, n9 = 9 * n * n;
    for (size_t p = 0; p < n; ++p) {
        orc_llg_diffusion_jacobian(w, m + 3 * p, srs[p], alphas[p]);
        for (size_t x = 0; x < 3; ++x)
            for (size_t y = 0; y < 3; ++y)
                for (size_t z = 0; z < 3; ++z)
                    jac[3 * p + z + (3 * p + y) * n3 + (3 * p + x) * n9] = w[z + 3 * y + 9 * x];
    }
}

/* ======================================================================
 * Fields, lib/field.cpp
 * ==================================================================== */

/* lib/field.cpp:37-40 */
double orc_field_sinusoidal(double t, double h, double f)
{
    return h * sin(2 * M_PI * f * t);
}

/* lib/field.cpp:51-54 */
double orc_field_square(double t, double h, double f)
{
    return h * ((int)(t * f * 2) % 2 ? -1 : 1);
}

/* lib/field.cpp:98-106 */
void orc_field_uniaxial(double *h, const double *m, const double *axis)
{
    double dot = m[0] * axis[0] + m[1] * axis[1] + m[2] * axis[2];
    h[0] = dot * axis[0];
    h[1] = dot * axis[1];
    h[2] = dot * axis[2];
}

/* lib/field.cpp:142-148 */
void orc_field_uniaxial_jacobian(double *jac, const double *axis)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            jac[i * 3 + j] = axis[i] * axis[j];
}

/* lib/field.cpp:116-132 */
void orc_field_multi_add_uniaxial(double *h, const double *m, const double *axes,
                                  const double *k_red, size_t n)
{
    for (size_t p = 0; p < n; ++p) {
        const double *e = axes + 3 * p, *s = m + 3 * p;
        double dot = s[0] * e[0] + s[1] * e[1] + s[2] * e[2];
        dot *= k_red[p];
        h[3 * p + 0] += dot * e[0];
        h[3 * p + 1] += dot * e[1];
        h[3 * p + 2] += dot * e[2];
    }
}

/* lib/field.cpp:159-174 */
void orc_field_multi_add_uniaxial_jacobian(double *jac, const double *axes,
                                           const double *k_red, size_t n)
{
    const size_t n3 = 3 * n;
    for (size_t i = 0; i < n3; ++i)
        for (size_t j = 0; j < n3; ++j)
            if (i / 3 == j / 3)
                jac[i * n3 + j] = k_red[i / 3] * axes[i] * axes[j];
}

/* lib/field.cpp:81-88 (the field function is evaluated once by the caller) */
void orc_field_multi_add_applied_z(double *h, double happ, size_t n)
{
    for (size_t p = 0; p < n; ++p) h[3 * p + 2] += happ;
}

/* lib/field.cpp:212-215 */
double orc_dipolar_prefactor(double ms, double k_av)
{
    return ORC_MU0 * ms * ms / 8.0 / M_PI / k_av;
}

/* lib/field.cpp:217-225 */
void orc_dipolar_add_p2p_term(double *out, double vj, double rij3, const double *mj,
                              const double *dist, double prefactor)
{
    double dot = mj[0] * dist[0] + mj[1] * dist[1] + mj[2] * dist[2];
    double t1 = vj / rij3;
    for (int i = 0; i < 3; ++i)
        out[i] += prefactor * t1 * (3 * dot * dist[i] - mj[i]);
}

/* lib/field.cpp:187-203 */
void orc_field_multi_add_dipolar(double *h, double ms, double k_av, const double *v_red,
                                 const double *m, const double *dists,
                                 const double *dist_cubes, size_t n)
{
    double pre = orc_dipolar_prefactor(ms, k_av);
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < i; ++j)
            orc_dipolar_add_p2p_term(h + 3 * i, v_red[j], dist_cubes[i * n + j], m + 3 * j,
                                     dists + i * n * 3 + j * 3, pre);
        for (size_t j = i + 1; j < n; ++j)
            orc_dipolar_add_p2p_term(h + 3 * i, v_red[j], dist_cubes[i * n + j], m + 3 * j,
                                     dists + i * n * 3 + j * 3, pre);
    }
}

/* ======================================================================
 * Distances, lib/distances.cpp
 * ==================================================================== */

/* lib/distances.cpp:20-35: out[i][j] = p_j - p_i */
void orc_pair_distance_vectors(double *out, const double *pts, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
            for (size_t k = 0; k < 3; ++k)
                out[(i * n + j) * 3 + k] = pts[3 * j + k] - pts[3 * i + k];
}

/* lib/distances.cpp:60-76 */
void orc_pair_distance_magnitudes(double *out, const double *v, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            const double *d = v + (i * n + j) * 3;
            out[i * n + j] = out[j * n + i] = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        }
}

/* lib/distances.cpp:101-113 (diagonal is 0/0 = NaN and never read) */
void orc_pair_unit_vectors(double *out, const double *v, const double *mags, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
            for (size_t k = 0; k < 3; ++k)
                out[(i * n + j) * 3 + k] = v[(i * n + j) * 3 + k] / mags[i * n + j];
}

/* ======================================================================
 * Integrators and Newton, lib/integrators.cpp, lib/optimisation.cpp
 * ==================================================================== */

/* lib/integrators.cpp:372-405 */
void orc_heun(double *next, double *drift, double *tdrift, double *diff, double *tdiff,
              const double *cur, const double *w, orc_sde_fn sde, void *ctx, size_t n,
              size_t wd, double t, double dt)
{
    sde(ctx, drift, diff, cur, t);
    for (size_t i = 0; i < n; ++i) {
        next[i] = cur[i] + dt * drift[i];
        for (size_t j = 0; j < wd; ++j)
            next[i] += diff[j + i * wd] * w[j] * sqrt(dt);
    }
    sde(ctx, tdrift, tdiff, next, t + dt);
    for (size_t i = 0; i < n; ++i) {
        next[i] = cur[i] + 0.5 * dt * (tdrift[i] + drift[i]);
        for (size_t j = 0; j < wd; ++j)
            next[i] += 0.5 * w[j] * sqrt(dt) * (tdiff[j + i * wd] + diff[j + i * wd]);
    }
}

/* BLAS dnrm2 (cblas_dnrm2 at lib/optimisation.cpp:98,142 and
 * lib/simulation.cpp:383): the Euclidean norm without overflow or underflow
 * for any finite double input.  The squares are summed in the x87 80-bit
 * format (64-bit significand, 15-bit exponent), whose range holds the square
 * of every double, and the root is rounded once more to double: the result
 * is within one ulp of the exact norm, as BLAS implementations are. */
double orc_dnrm2(int n, const double *x)
{
    long double s = 0.0L;
    for (int i = 0; i < n; ++i) s += (long double)x[i] * (long double)x[i];
    return (double)sqrtl(s);
}

/* LAPACK dgesv semantics for a row-major A (LAPACKE_dgesv_work, called at
 * lib/optimisation.cpp:134-137): LU with partial pivoting (first maximal
 * |a_ij| wins, as idamax), the column below the pivot scaled by its
 * reciprocal (dgetf2), then forward/back substitution.  Returns 0, or i > 0
 * when U(i,i) is exactly zero; b is untouched in that case. */
int orc_dgesv(int n, double *a, int *ipiv, double *b)
{
    int info = 0;
    for (int j = 0; j < n; ++j) {
        int p = j;
        double best = fabs(a[j * n + j]);
        for (int i = j + 1; i < n; ++i)
            if (fabs(a[i * n + j]) > best) { best = fabs(a[i * n + j]); p = i; }
        ipiv[j] = p;
        if (a[p * n + j] != 0.0) {
            if (p != j)
                for (int k = 0; k < n; ++k) {
                    double tmp = a[j * n + k]; a[j * n + k] = a[p * n + k]; a[p * n + k] = tmp;
                }
            if (fabs(a[j * n + j]) >= DBL_MIN) {
                double rcp = 1.0 / a[j * n + j];
                for (int i = j + 1; i < n; ++i) a[i * n + j] *= rcp;
            } else {
                for (int i = j + 1; i < n; ++i) a[i * n + j] /= a[j * n + j];
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int i = j + 1; i < n; ++i) {
            double l = a[i * n + j];
            for (int k = j + 1; k < n; ++k) a[i * n + k] -= l * a[j * n + k];
        }
    }
    if (info != 0) return info;
    for (int j = 0; j < n; ++j)
        if (ipiv[j] != j) { double t = b[j]; b[j] = b[ipiv[j]]; b[ipiv[j]] = t; }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < i; ++k) b[i] -= a[i * n + k] * b[k];
    for (int i = n - 1; i >= 0; --i) {
        for (int k = i + 1; k < n; ++k) b[i] -= a[i * n + k] * b[k];
        b[i] /= a[i * n + i];
    }
    return 0;
}

/* lib/optimisation.cpp:81-149.  Return codes as include/optimisation.hpp:
 * 0 = SUCCESS, 1 = MAX_ITERATIONS_ERR, 2 = LAPACK_ERR. */
int orc_newton_noinv(double *x_root, double *x_tmp, double *jac, int *ipiv, int *lapack_err,
                     orc_fj_fn fj, void *ctx, const double *x0, int dim, double eps,
                     size_t max_iter)
{
    for (int i = 0; i < dim; ++i) x_root[i] = x0[i];
    double tol = eps * orc_dnrm2(dim, x_root);
    double err = 2 * tol;
    long iter = (long)max_iter;
    while ((err > tol) && (iter-- > 0)) {
        for (int i = 0; i < dim; ++i) x_tmp[i] = x_root[i];
        fj(ctx, x_root, jac, x_tmp);
        for (int i = 0; i < dim; ++i) x_root[i] *= -1;
        *lapack_err = orc_dgesv(dim, jac, ipiv, x_root);
        if (*lapack_err != 0) return 2;
        err = orc_dnrm2(dim, x_root);
        for (int i = 0; i < dim; ++i) x_root[i] += x_tmp[i];
    }
    return iter == -1 ? 1 : 0;
}

typedef struct {
    orc_sdejac_fn sde;
    void *ctx;
    double *a, *b, *ad, *bd;
    const double *x0, *dwm;
    size_t n, w;
    double t, dt;
    int calls;
} orc_im_ctx;

/* the FJ closure of lib/integrators.cpp:617-637 */
static void orc_im_fj(void *vctx, double *fout, double *jac, const double *in)
{
    orc_im_ctx *c = (orc_im_ctx *)vctx;
    const size_t n = c->n, w = c->w;
    c->sde(c->ctx, c->a, c->b, c->ad, c->bd, in, c->t + c->dt / 2, c->t);
    c->calls++;
    for (size_t i = 0; i < n; ++i) {
        fout[i] = in[i] - 0.5 * c->a[i] * c->dt - c->x0[i];
        for (size_t j = 0; j < w; ++j) fout[i] -= 0.5 * c->b[i * w + j] * c->dwm[j];
    }
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            jac[i * n + j] = (double)(i == j) - 0.5 * c->ad[i * n + j];
            for (size_t k = 0; k < w; ++k)
                jac[i * n + j] -= 0.5 * c->bd[i * w * n + k * n + j] * c->dwm[k];
        }
}

/* lib/integrators.cpp:576-651 */
int orc_implicit_midpoint(double *x, double *dwm, double *a_work, double *b_work,
                          double *adash_work, double *bdash_work, double *x_guess,
                          double *x_opt_tmp, double *x_opt_jac, int *ipiv,
                          const double *x0, const double *dw, orc_sdejac_fn sde, void *ctx,
                          size_t n, size_t wd, double t, double dt, double eps,
                          size_t max_iter, int *newton_iters)
{
    const double k = 1000;
    const double ah = sqrt(2 * k * fabs(log(dt)));
    for (size_t i = 0; i < wd; ++i)
        dwm[i] = fmax(-ah, fmin(ah, dw[i])) * sqrt(dt);

    sde(ctx, a_work, b_work, adash_work, bdash_work, x0, t, t);
    for (size_t i = 0; i < n; ++i) {
        x_guess[i] = a_work[i] * dt;
        for (size_t j = 0; j < wd; ++j) x_guess[i] += b_work[i * wd + j] * dwm[j];
    }
    for (size_t i = 0; i < n; ++i) x_guess[i] = (x_guess[i] + x0[i]) / 2;

    orc_im_ctx c = {sde, ctx, a_work, b_work, adash_work, bdash_work, x0, dwm, n, wd,
                    t, dt, 0};
    int err_code = 0;
    int flag = orc_newton_noinv(x, x_opt_tmp, x_opt_jac, ipiv, &err_code, orc_im_fj, &c,
                                x_guess, (int)n, eps, max_iter);
    if (newton_iters) *newton_iters = c.calls;
    for (size_t i = 0; i < n; ++i) x[i] = 2 * x[i] - x0[i];
    return flag + err_code;
}

/* lib/trap.cpp:16-29 */
double orc_trapezoidal(const double *x, const double *y, size_t n)
{
    double sum = 0.0;
    for (size_t i = 1; i < n; ++i) sum += (x[i] - x[i - 1]) * (y[i] + y[i - 1]);
    sum /= 2.0;
    return sum;
}

/* ======================================================================
 * Random-number sources used by full dynamics
 * ==================================================================== */
void orc_rng_init_mt(orc_rng *r, uint64_t seed, double stddev)
{
    memset(r, 0, sizeof(*r));
    r->kind = ORC_RNG_MT;
    orc_mtnorm_init(&r->mt, seed, stddev);
}

void orc_rng_init_array(orc_rng *r, const double *arr, size_t len, size_t stride)
{
    memset(r, 0, sizeof(*r));
    r->kind = ORC_RNG_ARRAY;
    r->arr = arr;
    r->arr_len = len;
    r->arr_stride = stride ? stride : 1;
}

void orc_rng_init_philox(orc_rng *r, uint64_t seed, uint32_t tag, uint32_t n_particles)
{
    memset(r, 0, sizeof(*r));
    r->kind = ORC_RNG_PHILOX;
    r->seed = seed;
    r->tag = tag;
    r->n_particles = n_particles;
}

/* RngMtNorm::get (lib/rng.cpp:24), RngArray::get (lib/rng.cpp:84-94; the
 * out_of_range throw becomes the `exhausted` flag and a NaN draw), and the
 * Philox stream drawn in the reference's order: 3N per step, particle-major
 * (lib/simulation.cpp:358-359). */
double orc_rng_get(orc_rng *r)
{
    switch (r->kind) {
    case ORC_RNG_MT:
        return orc_mtnorm_get(&r->mt);
    case ORC_RNG_ARRAY:
        if (r->arr_i < r->arr_len) {
            size_t i = r->arr_i;
            r->arr_i += r->arr_stride;
            return r->arr[i];
        }
        r->exhausted = 1;
        return NAN;
    default: {
        uint32_t p = r->idx / 3, d = r->idx % 3;
        double z;
        orc_wiener_normals(r->seed, p, r->tag, 3 * r->step + d, 1, &z);
        if (++r->idx == 3 * r->n_particles) { r->idx = 0; r->step++; }
        return z;
    }
    }
}

/* ======================================================================
 * Full dynamics, lib/simulation.cpp
 * ==================================================================== */
typedef struct {
    size_t n;
    const double *axes, *k_red, *v_red, *dists, *dist_cubes, *alphas, *srs;
    int field_shape, interactions;
    double happ, fapp, ms, k_av;
    double *heff, *heffjac;
} orc_sim_ctx;

/* the field-function lambdas of lib/simulation.cpp:552-573 */
static double orc_applied(const orc_sim_ctx *c, double t)
{
    switch (c->field_shape) {
    case ORC_FIELD_SINE: return orc_field_sinusoidal(t, c->happ, c->fapp);
    case ORC_FIELD_SQUARE: return orc_field_square(t, c->happ, c->fapp);
    default: return c->happ;
    }
}

/* heff_func, lib/simulation.cpp:271-290 */
static void orc_heff(const orc_sim_ctx *c, double *h, const double *m, double t)
{
    for (size_t i = 0; i < 3 * c->n; ++i) h[i] = 0.0;
    orc_field_multi_add_uniaxial(h, m, c->axes, c->k_red, c->n);
    orc_field_multi_add_applied_z(h, orc_applied(c, t), c->n);
    if (c->interactions)
        orc_field_multi_add_dipolar(h, c->ms, c->k_av, c->v_red, m, c->dists, c->dist_cubes,
                                    c->n);
}

/* sde_no_jacs, lib/simulation.cpp:328-336 */
static void orc_sde_nojac(void *vctx, double *drift, double *diff, const double *m, double t)
{
    orc_sim_ctx *c = (orc_sim_ctx *)vctx;
    orc_heff(c, c->heff, m, t);
    orc_llg_multi_drift(drift, m, c->alphas, c->heff, c->n);
    orc_llg_multi_diffusion(diff, m, c->srs, c->alphas, c->n);
}

/* sde, lib/simulation.cpp:306-325 -> lib/llg.cpp:453-481; the field
 * Jacobian ignores interactions (lib/simulation.cpp:296-303) */
static void orc_sde_jac(void *vctx, double *drift, double *diff, double *jd, double *jb,
                        const double *m, double ta, double tb)
{
    (void)tb;
    orc_sim_ctx *c = (orc_sim_ctx *)vctx;
    const size_t n3 = 3 * c->n;
    orc_heff(c, c->heff, m, ta);
    orc_llg_multi_drift(drift, m, c->alphas, c->heff, c->n);
    orc_llg_multi_diffusion(diff, m, c->srs, c->alphas, c->n);
    for (size_t i = 0; i < n3 * n3; ++i) c->heffjac[i] = 0.0;
    orc_field_multi_add_uniaxial_jacobian(c->heffjac, c->axes, c->k_red, c->n);
    orc_llg_multi_drift_quasijacobian(jd, m, c->alphas, c->heff, c->heffjac, c->n);
    orc_llg_multi_diffusion_jacobian(jb, m, c->srs, c->alphas, c->n);
}

int64_t orc_sample_schedule(double dt, double end_time, size_t max_samples, int64_t *k)
{
    if (max_samples < 1) return -1;
    const double ts = end_time / (double)(max_samples - 1);
    int64_t step = 0;
    double t = 0;
    k[0] = 0;
    for (size_t s = 1; s < max_samples; ++s) {
        while (t <= (double)s * ts) {
            step++;
            t = (double)step * dt;
        }
        k[s] = step - 1;
    }
    return max_samples > 1 ? step : 0;
}

int orc_full_dynamics_reduced(
    const double *sr, const double *k_red, const double *v_red, const double *axes,
    const double *m0, const double *unit_dists, const double *red_dists, size_t n,
    int field_shape, double happ, double fapp, double average_anisotropy,
    double average_volume, double damping, double ms, double dt, double end_time,
    orc_rng *rng, int renorm, int interactions, int use_implicit, double eps,
    int max_samples, double *time, double *field, double *mx, double *my, double *mz,
    orc_stats *stats)
{
    (void)average_volume;
    /* max_samples == -1 never leaves the step loop in the reference
     * (lib/simulation.cpp:345,355 keep t == sample*Ts); rejected here. */
    if (max_samples < 1 || n == 0) return -1;
    const size_t n3 = 3 * n, n33 = n3 * n3;
    const size_t ns = (size_t)max_samples;
    const double ts = end_time / (double)(ns - 1);
    const size_t max_iter = 1000; /* lib/simulation.cpp:204 */

    double *a_work = calloc(n3, sizeof(double));
    double *b_work = calloc(n33, sizeof(double));
    double *adash = calloc(n33, sizeof(double));
    /* The reference allocates the (3N)^3 tensor for Heun too
     * (lib/simulation.cpp:194-195) but Heun only uses its first (3N)^2
     * entries as the trial diffusion matrix. */
    double *bdash = calloc(use_implicit ? n33 * n3 : n33, sizeof(double));
    double *dwm = calloc(n3, sizeof(double));
    double *x_guess = calloc(n3, sizeof(double)), *x_tmp = calloc(n3, sizeof(double));
    double *x_jac = calloc(n33, sizeof(double));
    int *ipiv = calloc(n3, sizeof(int));
    double *wiener = calloc(n3, sizeof(double));
    double *heff = calloc(n3, sizeof(double)), *heffjac = calloc(n33, sizeof(double));
    double *pstate = calloc(n3, sizeof(double)), *nstate = calloc(n3, sizeof(double));
    double *cubes = calloc(n * n, sizeof(double));
    double *alphas = calloc(n, sizeof(double));
    if (!a_work || !b_work || !adash || !bdash || !dwm || !x_guess || !x_tmp || !x_jac ||
        !ipiv || !wiener || !heff || !heffjac || !pstate || !nstate || !cubes || !alphas)
        return -2;

    for (size_t p = 0; p < n; ++p) alphas[p] = damping;
    for (size_t i = 0; i < n * n; ++i) cubes[i] = pow(red_dists[i], 3);

    orc_sim_ctx ctx = {n, axes, k_red, v_red, unit_dists, cubes, alphas, sr, field_shape,
                       interactions, happ, fapp, ms, average_anisotropy, heff, heffjac};

    time[0] = 0;
    field[0] = orc_applied(&ctx, 0);
    for (size_t p = 0; p < n; ++p) {
        mx[p * ns] = m0[3 * p];
        my[p * ns] = m0[3 * p + 1];
        mz[p * ns] = m0[3 * p + 2];
    }
    for (size_t i = 0; i < n3; ++i) nstate[i] = m0[i];

    unsigned int step = 0;
    double t = 0;
    int64_t newton_total = 0, newton_fail = 0;
    for (size_t sample = 1; sample < ns; ++sample) {
        while (t <= (double)sample * ts) {
            for (size_t i = 0; i < n3; ++i) pstate[i] = nstate[i];
            step++;
            t = (double)step * dt;
            for (size_t i = 0; i < n3; ++i) wiener[i] = orc_rng_get(rng);
            if (use_implicit) {
                int it = 0;
                int code = orc_implicit_midpoint(nstate, dwm, a_work, b_work, adash, bdash,
                                                 x_guess, x_tmp, x_jac, ipiv, pstate, wiener,
                                                 orc_sde_jac, &ctx, n3, n3, t, dt, eps,
                                                 max_iter, &it);
                newton_total += it;
                if (code != 0) newton_fail++;
            } else {
                orc_heun(nstate, a_work, adash, b_work, bdash, pstate, wiener, orc_sde_nojac,
                         &ctx, n3, n3, t, dt);
            }
            if (renorm) {
                for (size_t off = 0; off < n3; off += 3) {
                    double norm = orc_dnrm2(3, nstate + off);
                    for (int i = 0; i < 3; ++i) nstate[off + i] = nstate[off + i] / norm;
                }
            }
        }
        time[sample] = (double)sample * ts;
        field[sample] = orc_applied(&ctx, (double)sample * ts);
        for (size_t p = 0; p < n; ++p) {
            mx[p * ns + sample] = pstate[3 * p];
            my[p * ns + sample] = pstate[3 * p + 1];
            mz[p * ns + sample] = pstate[3 * p + 2];
        }
    }
    if (stats) {
        stats->n_steps = step;
        stats->newton_iters = newton_total;
        stats->newton_failures = newton_fail;
        stats->rng_exhausted = rng->kind == ORC_RNG_ARRAY ? rng->exhausted : 0;
    }
    free(a_work); free(b_work); free(adash); free(bdash); free(dwm); free(x_guess);
    free(x_tmp); free(x_jac); free(ipiv); free(wiener); free(heff); free(heffjac);
    free(pstate); free(nstate); free(cubes); free(alphas);
    return 0;
}

int orc_full_dynamics(
    const double *radius, const double *anisotropy, const double *axes, const double *m0,
    const double *location, size_t n, double magnetisation, double damping,
    double temperature, int renorm, int interactions, int use_implicit, double eps,
    double time_step, double end_time, size_t max_samples, int64_t seed,
    int field_shape, double field_amplitude, double field_frequency,
    int rng_kind, const double *noise, size_t noise_len,
    double *time, double *field, double *mx, double *my, double *mz, orc_stats *stats)
{
    if (n == 0 || max_samples < 1) return -1;
    if (field_shape != ORC_FIELD_SINE && field_shape != ORC_FIELD_SQUARE &&
        field_shape != ORC_FIELD_CONSTANT)
        return -3; /* invalid_argument, lib/simulation.cpp:570-572 */
    double *volume = malloc(n * sizeof(double));
    double *v_red = malloc(n * sizeof(double));
    double *k_red = malloc(n * sizeof(double));
    double *sr = malloc(n * sizeof(double));
    double *vecs = malloc(n * n * 3 * sizeof(double));
    double *mags = malloc(n * n * sizeof(double));
    double *units = malloc(n * n * 3 * sizeof(double));

    /* lib/simulation.cpp:501-510 */
    for (size_t p = 0; p < n; ++p)
        volume[p] = 4.0 / 3.0 * M_PI * radius[p] * radius[p] * radius[p];
    double vsum = 0.0;
    for (size_t p = 0; p < n; ++p) vsum += volume[p];
    const double average_volume = vsum / (double)n;
    for (size_t p = 0; p < n; ++p) v_red[p] = volume[p] / average_volume;
    /* lib/simulation.cpp:519-524 */
    double ksum = 0.0;
    for (size_t p = 0; p < n; ++p) ksum += anisotropy[p];
    const double average_anisotropy = ksum / (double)n;
    for (size_t p = 0; p < n; ++p) k_red[p] = anisotropy[p] / average_anisotropy;
    /* lib/simulation.cpp:527-533 */
    const double anisotropy_field = 2 * average_anisotropy / ORC_MU0 / magnetisation;
    const double time_factor = ORC_GYROMAG * ORC_MU0 * anisotropy_field / (1 + damping * damping);
    const double rdt = time_step * time_factor;
    const double rend = end_time * time_factor;
    /* lib/simulation.cpp:536-541 */
    for (size_t p = 0; p < n; ++p)
        sr[p] = sqrt(damping * ORC_KB * temperature / (average_anisotropy * volume[p]) /
                     (1 + damping * damping));
    /* lib/simulation.cpp:548-549 */
    const double happ = field_amplitude / anisotropy_field;
    const double fapp = field_frequency / time_factor;
    /* lib/simulation.cpp:577-583 */
    orc_pair_distance_vectors(vecs, location, n);
    orc_pair_distance_magnitudes(mags, vecs, n);
    orc_pair_unit_vectors(units, vecs, mags, n);
    const double vroot = pow(average_volume, 1. / 3);
    for (size_t i = 0; i < n * n; ++i) mags[i] /= vroot;

    orc_rng rng;
    if (rng_kind == ORC_RNG_MT)
        orc_rng_init_mt(&rng, (uint64_t)seed, 1.0); /* lib/simulation.cpp:586 */
    else if (rng_kind == ORC_RNG_ARRAY)
        orc_rng_init_array(&rng, noise, noise_len, 1);
    else
        orc_rng_init_philox(&rng, (uint64_t)seed, 0, (uint32_t)n);

    int rc = orc_full_dynamics_reduced(sr, k_red, v_red, axes, m0, units, mags, n, field_shape,
                                       happ, fapp, average_anisotropy, average_volume, damping,
                                       magnetisation, rdt, rend, &rng, renorm, interactions,
                                       use_implicit, eps, (int)max_samples, time, field, mx, my,
                                       mz, stats);
    if (rc == 0) {
        /* lib/simulation.cpp:610-621 */
        for (size_t i = 0; i < max_samples; ++i) {
            time[i] /= time_factor;
            field[i] *= anisotropy_field;
        }
        for (size_t i = 0; i < n * max_samples; ++i) {
            mx[i] *= magnetisation;
            my[i] *= magnetisation;
            mz[i] *= magnetisation;
        }
    }
    free(volume); free(v_red); free(k_red); free(sr); free(vecs); free(mags); free(units);
    return rc;
}

/* ======================================================================
 * Pathwise sensitivities of the Heun path (see sllg_oracle.h).  The
 * tangent differentiates each operation of the step directly (forward
 * mode, product rule), independently of any transposed form.
 * ==================================================================== */
double orc_path_phase(double fdt, int64_t k)
{
    double c = (double)k * fdt;
    c -= floor(c);
    return sin(2.0 * M_PI * c);
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* L(m, v) = -m x v + alpha (v |m|^2 - m (m.v)) */
static void path_L(const double *m, const double *v, double alpha, double *out)
{
    double mm = dot3(m, m), mv = dot3(m, v);
    out[0] = -(m[1] * v[2] - m[2] * v[1]) + alpha * (v[0] * mm - m[0] * mv);
    out[1] = -(m[2] * v[0] - m[0] * v[2]) + alpha * (v[1] * mm - m[1] * mv);
    out[2] = -(m[0] * v[1] - m[1] * v[0]) + alpha * (v[2] * mm - m[2] * mv);
}

/* directional derivative of L along (dm, dv, dalpha) */
static void path_dL(const double *m, const double *v, double alpha, const double *dm,
                    const double *dv, double dalpha, double *out)
{
    double mm = dot3(m, m), mv = dot3(m, v);
    double dmm = 2.0 * dot3(m, dm), dmv = dot3(dm, v) + dot3(m, dv);
    double dc[3];
    /* d(-m x v) = -dm x v - m x dv */
    dc[0] = -(dm[1] * v[2] - dm[2] * v[1]) - (m[1] * dv[2] - m[2] * dv[1]);
    dc[1] = -(dm[2] * v[0] - dm[0] * v[2]) - (m[2] * dv[0] - m[0] * dv[2]);
    dc[2] = -(dm[0] * v[1] - dm[1] * v[0]) - (m[0] * dv[1] - m[1] * dv[0]);
    for (int i = 0; i < 3; ++i) {
        double g = v[i] * mm - m[i] * mv;
        double dg = dv[i] * mm + v[i] * dmm - dm[i] * mv - m[i] * dmv;
        out[i] = dc[i] + dalpha * g + alpha * dg;
    }
}

static void path_v(const double *m, const double *e, const double *w, double dt, double sig,
                   double happ, double s, double *v)
{
    double me = dot3(m, e);
    for (int i = 0; i < 3; ++i) v[i] = dt * (me * e[i] + (i == 2 ? happ * s : 0.0)) + sig * w[i];
}

static void path_dv(const double *m, const double *dm, const double *e, const double *w,
                    double dt, double happ, double s, const double *dq, double *dv)
{
    double me = dot3(m, e), dme = dot3(dm, e);
    for (int i = 0; i < 3; ++i) {
        double h = me * e[i] + (i == 2 ? happ * s : 0.0);
        double dh = dme * e[i] + (i == 2 ? dq[3] * s : 0.0);
        dv[i] = dq[0] * h + dt * dh + dq[1] * w[i];
    }
}

void orc_path_step(double m[3], const double e[3], const double w[3], double dt, double sig,
                   double alpha, double happ, double s1, double s2)
{
    double v1[3], a1[3], mt[3], v2[3], a2[3];
    path_v(m, e, w, dt, sig, happ, s1, v1);
    path_L(m, v1, alpha, a1);
    for (int i = 0; i < 3; ++i) mt[i] = m[i] + a1[i];
    path_v(mt, e, w, dt, sig, happ, s2, v2);
    path_L(mt, v2, alpha, a2);
    for (int i = 0; i < 3; ++i) m[i] = m[i] + 0.5 * (a1[i] + a2[i]);
}

void orc_path_run(const orc_path_problem *pr, double *mz_sum, double *mz_final_sum,
                  double *final_state)
{
    for (size_t s = 0; s < pr->n_samples; ++s) mz_sum[s] = 0.0;
    *mz_final_sum = 0.0;
    for (size_t p = 0; p < pr->n_particles; ++p) {
        double m[3] = {pr->m0[3 * p], pr->m0[3 * p + 1], pr->m0[3 * p + 2]};
        const double *e = pr->axis + 3 * p;
        size_t s = 0;
        for (int64_t n = 0;; ++n) {
            while (s < pr->n_samples && pr->sample_steps[s] == n) mz_sum[s++] += m[2];
            if (n == pr->n_steps) break;
            double w[3];
            orc_wiener_normals(pr->seed[p], 0, pr->tag, 3 * (uint64_t)n, 3, w);
            orc_path_step(m, e, w, pr->dt[p], pr->sig[p], pr->alpha[p], pr->happ[p],
                          orc_path_phase(pr->fdt, n + 1), orc_path_phase(pr->fdt, n + 2));
        }
        *mz_final_sum += m[2];
        if (final_state)
            for (int i = 0; i < 3; ++i) final_state[3 * p + i] = m[i];
    }
}

void orc_path_tangent(const orc_path_problem *pr, size_t n_dirs, const double *dq,
                      double *dmz_sum, double *dmz_final_sum)
{
    for (size_t i = 0; i < pr->n_samples * n_dirs; ++i) dmz_sum[i] = 0.0;
    for (size_t d = 0; d < n_dirs; ++d) dmz_final_sum[d] = 0.0;
    double *tm = (double *)calloc(3 * n_dirs, sizeof(double));
    for (size_t p = 0; p < pr->n_particles; ++p) {
        double m[3] = {pr->m0[3 * p], pr->m0[3 * p + 1], pr->m0[3 * p + 2]};
        const double *e = pr->axis + 3 * p;
        const double dt = pr->dt[p], sig = pr->sig[p], al = pr->alpha[p], ha = pr->happ[p];
        memset(tm, 0, 3 * n_dirs * sizeof(double));
        size_t s = 0;
        for (int64_t n = 0;; ++n) {
            while (s < pr->n_samples && pr->sample_steps[s] == n) {
                for (size_t d = 0; d < n_dirs; ++d) dmz_sum[s * n_dirs + d] += tm[3 * d + 2];
                ++s;
            }
            if (n == pr->n_steps) break;
            double w[3];
            orc_wiener_normals(pr->seed[p], 0, pr->tag, 3 * (uint64_t)n, 3, w);
            const double s1 = orc_path_phase(pr->fdt, n + 1), s2 = orc_path_phase(pr->fdt, n + 2);
            double v1[3], a1[3], mt[3], v2[3], a2[3];
            path_v(m, e, w, dt, sig, ha, s1, v1);
            path_L(m, v1, al, a1);
            for (int i = 0; i < 3; ++i) mt[i] = m[i] + a1[i];
            path_v(mt, e, w, dt, sig, ha, s2, v2);
            path_L(mt, v2, al, a2);
            for (size_t d = 0; d < n_dirs; ++d) {
                const double *q = dq + (p * n_dirs + d) * 4;
                double *dm = tm + 3 * d;
                double dv1[3], da1[3], dmt[3], dv2[3], da2[3];
                path_dv(m, dm, e, w, dt, ha, s1, q, dv1);
                path_dL(m, v1, al, dm, dv1, q[2], da1);
                for (int i = 0; i < 3; ++i) dmt[i] = dm[i] + da1[i];
                path_dv(mt, dmt, e, w, dt, ha, s2, q, dv2);
                path_dL(mt, v2, al, dmt, dv2, q[2], da2);
                for (int i = 0; i < 3; ++i) dm[i] = dm[i] + 0.5 * (da1[i] + da2[i]);
            }
            for (int i = 0; i < 3; ++i) m[i] = m[i] + 0.5 * (a1[i] + a2[i]);
        }
        for (size_t d = 0; d < n_dirs; ++d) dmz_final_sum[d] += tm[3 * d + 2];
    }
    free(tm);
}

/* ======================================================================
 * Discrete orientation model
 * ==================================================================== */
double orc_field_square_fourier(double t, double h, double f, size_t n_components)
{
    double field = 0.0;
    for (size_t k = 1; k < n_components + 1; ++k) {
        double odd = (double)(2 * k - 1);
        field += sin(2 * M_PI * odd * f * t) / odd;
    }
    field *= 4 / M_PI * h;
    return field;
}

void orc_dom_transition_matrix(double *W, double k, double v, double T, double h, double ms,
                               double alpha)
{
    double sigma = k * v / ORC_KB / T;
    double taun = v * ms * (1 + alpha * alpha) / 2.0 / ORC_GYROMAG / alpha / ORC_KB / T;
    double e1 = sigma * (1 - h) * (1 - h);
    double e2 = sigma * (1 + h) * (1 + h);
    double pre = taun * sqrt(M_PI) / pow(sigma, 1.5) / (1 - h * h);
    double r1 = 1.0 / pre * (1 - h) * exp(-e1);
    double r2 = 1.0 / pre * (1 + h) * exp(-e2);
    W[0] = -r2;
    W[1] = r1;
    W[2] = r2;
    W[3] = -r1;
}

void orc_master_equation(double *derivs, const double *W, const double *x, size_t dim)
{
    for (size_t i = 0; i < dim; ++i) {
        double acc = 0.0;
        for (size_t j = 0; j < dim; ++j) acc += W[i * dim + j] * x[j];
        derivs[i] = acc;
    }
}

/* Cash-Karp coefficients, include/integrators.hpp:128-160 */
static const double CK_C11 = 0.2, CK_C21 = 3.0 / 40.0, CK_C22 = 9.0 / 40.0, CK_C31 = 3.0 / 10.0,
                    CK_C32 = -9.0 / 10.0, CK_C33 = 6.0 / 5.0, CK_C41 = -11.0 / 54.0,
                    CK_C42 = 2.5, CK_C43 = -70.0 / 27.0, CK_C44 = 35.0 / 27.0,
                    CK_C51 = 1631.0 / 55296.0, CK_C52 = 175.0 / 512.0, CK_C53 = 575.0 / 13824.0,
                    CK_C54 = 44275.0 / 110592.0, CK_C55 = 253.0 / 4096.0;
static const double CK_H1 = 0.2, CK_H2 = 0.3, CK_H3 = 0.6, CK_H4 = 1.0, CK_H5 = 7.0 / 8.0;
static const double CK_X11 = 37.0 / 378.0, CK_X13 = 250.0 / 621.0, CK_X14 = 125.0 / 594.0,
                    CK_X16 = 512.0 / 1771.0, CK_X21 = 2825.0 / 27648.0,
                    CK_X23 = 18575.0 / 48384.0, CK_X24 = 13525.0 / 55296.0,
                    CK_X25 = 277.0 / 14336.0, CK_X26 = 0.25;

void orc_rk45(double *next_state, double *temp_state, double *k1, double *k2, double *k3,
              double *k4, double *k5, double *k6, double *h_ptr, double *t_ptr,
              const double *x, orc_ode_fn derivs, void *ctx, size_t n, double tol)
{
    int ok = 0;
    double err = 0.0, h = *h_ptr, t = *t_ptr;
    while (!ok) {
        derivs(ctx, k1, x, t);
        for (size_t i = 0; i < n; ++i) next_state[i] = k1[i] * h * CK_C11 + x[i];
        derivs(ctx, k2, next_state, t + h * CK_H1);
        for (size_t i = 0; i < n; ++i) next_state[i] = x[i] + h * (CK_C21 * k1[i] + CK_C22 * k2[i]);
        derivs(ctx, k3, next_state, t + h * CK_H2);
        for (size_t i = 0; i < n; ++i)
            next_state[i] = x[i] + h * (CK_C31 * k1[i] + CK_C32 * k2[i] + CK_C33 * k3[i]);
        derivs(ctx, k4, next_state, t + h * CK_H3);
        for (size_t i = 0; i < n; ++i)
            next_state[i] = x[i] + h * (CK_C41 * k1[i] + CK_C42 * k2[i] + CK_C43 * k3[i] +
                                        CK_C44 * k4[i]);
        derivs(ctx, k5, next_state, t + h * CK_H4);
        for (size_t i = 0; i < n; ++i)
            next_state[i] = x[i] + h * (CK_C51 * k1[i] + CK_C52 * k2[i] + CK_C53 * k3[i] +
                                        CK_C54 * k4[i] + CK_C55 * k5[i]);
        derivs(ctx, k6, next_state, t + h * CK_H5);
        for (size_t i = 0; i < n; ++i)
            temp_state[i] = x[i] + h * (CK_X11 * k1[i] + CK_X13 * k3[i] + CK_X14 * k4[i] +
                                        CK_X16 * k6[i]);
        for (size_t i = 0; i < n; ++i)
            next_state[i] = x[i] + h * (CK_X21 * k1[i] + CK_X23 * k3[i] + CK_X24 * k4[i] +
                                        CK_X25 * k5[i] + CK_X26 * k6[i]);
        err = 0.0;
        double mag = 0.0;
        for (size_t i = 0; i < n; ++i) mag += x[i] * x[i];
        mag = pow(mag, 0.5);
        for (size_t i = 0; i < n; ++i) err += pow(fabs(temp_state[i] - next_state[i]), 2);
        err = pow(err, 0.5);
        err /= ((double)n * tol * (1 + mag));
        if (err < 1.0) {
            ok = 1;
        } else {
            double f = 0.84 * pow(err, -0.2);
            f = fabs(f) < 0.1 ? 0.1 : f;
            h *= f;
        }
    }
    *t_ptr = t + h;
    double f = err == 0.0 ? 5.0 : 0.84 * pow(err, -0.2);
    f = f > 5 ? 5.0 : f;
    *h_ptr = f * h;
}

typedef struct {
    double k, v, T, ms, alpha, h, f;
    size_t ncomp;
    int shape;
} orc_dom_ctx;

static double orc_dom_field(const orc_dom_ctx *c, double t)
{
    switch (c->shape) {
    case ORC_FIELD_SINE: return orc_field_sinusoidal(t, c->h, c->f);
    case ORC_FIELD_SQUARE: return orc_field_square(t, c->h, c->f);
    case ORC_FIELD_SQUARE_F: return orc_field_square_fourier(t, c->h, c->f, c->ncomp);
    default: return c->h;
    }
}

/* dom::master_equation_with_update (lib/dom.cpp:80-93) */
static void orc_dom_master(void *ctx, double *dx, const double *x, double t)
{
    const orc_dom_ctx *c = (const orc_dom_ctx *)ctx;
    double W[4];
    orc_dom_transition_matrix(W, c->k, c->v, c->T, orc_dom_field(c, t), c->ms, c->alpha);
    orc_master_equation(dx, W, x, 2);
}

void orc_dom_ensemble_dynamics(double volume, double anisotropy, double temperature,
                               double magnetisation, double alpha, int field_shape,
                               double h_reduced, double frequency, size_t n_components,
                               const double initial_probs[2], double time_step, double end_time,
                               size_t max_samples, double *time, double *field, double *mz)
{
    orc_dom_ctx c = {anisotropy, volume, temperature, magnetisation, alpha, h_reduced, frequency,
                     n_components, field_shape};
    double last[2], next[2] = {initial_probs[0], initial_probs[1]};
    double k1[2], k2[2], k3[2], k4[2], k5[2], k6[2], tmp[2];
    const size_t S = max_samples;
    const double ts = end_time / (double)(S - 1);
    mz[0] = next[0] - next[1];
    time[0] = 0;
    field[0] = orc_dom_field(&c, 0);
    double t = 0, t_last = 0;
    const double max_dt = end_time / 1000.0;
    double dt = 0.01 * time_step;
    for (size_t s = 1; s < S; ++s) {
        while (t <= (double)s * ts) {
            last[0] = next[0];
            last[1] = next[1];
            t_last = t;
            orc_rk45(next, tmp, k1, k2, k3, k4, k5, k6, &dt, &t, last, orc_dom_master, &c, 2,
                     time_step);
            dt = dt > max_dt ? max_dt : dt;
        }
        double mz_last = last[0] - last[1], mz_next = next[0] - next[1];
        double ts_s = (double)s * ts;
        double beta = (mz_next - mz_last) / (t - t_last);
        mz[s] = mz_last + beta * (ts_s - t_last);
        time[s] = ts_s;
        field[s] = orc_dom_field(&c, ts_s);
    }
}
