// ref_capi.cpp -- C entry points over owlas/magpy's own building blocks.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's lib/llg.cpp, lib/field.cpp, lib/rng.cpp, lib/distances.cpp
// and lib/trap.cpp, read where they lie, into oracle/_ref/libmagpy_ref.so
// (include/io.hpp's write_array is header-only and comes along).
// The tests use it to check the oracle's restatement of those functions.
#include "llg.hpp"
#include "field.hpp"
#include "rng.hpp"
#include "distances.hpp"
#include "trap.hpp"
// io.hpp uses std::runtime_error and FILE without including their headers
#include <cstdio>
#include <stdexcept>
#include "io.hpp"
#include <array>
#include <vector>

extern "C" {

void ref_llg_drift(double *d, const double *m, double alpha, const double *h)
{
    llg::drift(d, m, 0.0, alpha, h);
}

void ref_llg_diffusion(double *d, const double *m, double sr, double alpha)
{
    llg::diffusion(d, m, 0.0, sr, alpha);
}

void ref_llg_drift_jacobian(double *j, const double *m, double a, const double *h,
                            const double *hj)
{
    llg::drift_jacobian(j, m, 0.0, a, h, hj);
}

void ref_llg_diffusion_jacobian(double *j, const double *m, double sr, double alpha)
{
    llg::diffusion_jacobian(j, m, 0.0, sr, alpha);
}

void ref_llg_multi_diffusion(double *d, const double *m, const double *srs,
                             const double *alphas, size_t n)
{
    llg::multi_diffusion(d, m, srs, alphas, n);
}

void ref_llg_multi_drift_quasijacobian(double *jac, const double *m, const double *alphas,
                                       const double *h, const double *hj, size_t n)
{
    llg::multi_drift_quasijacobian(jac, m, alphas, h, hj, n);
}

void ref_llg_multi_diffusion_jacobian(double *jac, const double *m, const double *srs,
                                      const double *alphas, size_t n)
{
    llg::multi_diffusion_jacobian(jac, m, srs, alphas, n);
}

double ref_field_sinusoidal(double t, double h, double f) { return field::sinusoidal(t, h, f); }
double ref_field_square(double t, double h, double f) { return field::square(t, h, f); }

void ref_field_uniaxial(double *h, const double *m, const double *axis)
{
    field::uniaxial_anisotropy(h, m, axis);
}

void ref_field_multi_add_uniaxial(double *h, const double *m, const double *axes,
                                  const double *k_red, size_t n)
{
    field::multi_add_uniaxial_anisotropy(h, m, axes, k_red, n);
}

void ref_field_multi_add_uniaxial_jacobian(double *jac, const double *axes,
                                           const double *k_red, size_t n)
{
    field::multi_add_uniaxial_anisotropy_jacobian(jac, axes, k_red, n);
}

void ref_field_multi_add_dipolar(double *h, double ms, double k_av, const double *v_red,
                                 const double *m, const double *dists,
                                 const double *dist_cubes, size_t n)
{
    field::multi_add_dipolar(h, ms, k_av, v_red, m, dists, dist_cubes, n);
}

double ref_dipolar_prefactor(double ms, double k_av) { return field::dipolar_prefactor(ms, k_av); }

// n draws of RngMtNorm(seed, std)
void ref_rng_mtnorm(unsigned long seed, double stddev, double *out, size_t n)
{
    RngMtNorm rng(seed, stddev);
    for (size_t i = 0; i < n; ++i) out[i] = rng.get();
}

double ref_trapezoidal(double *x, double *y, size_t n) { return trap::trapezoidal(x, y, n); }

// unit vectors [n*n*3] and magnitudes [n*n] of points [n*3]
void ref_distances(const double *pts, size_t n, double *units, double *mags)
{
    std::vector<std::array<double, 3>> p(n);
    for (size_t i = 0; i < n; ++i)
        for (size_t k = 0; k < 3; ++k) p[i][k] = pts[3 * i + k];
    auto vecs = distances::pair_wise_distance_vectors(p);
    auto m = distances::pair_wise_distance_magnitude(vecs);
    auto u = distances::pair_wise_distance_unit_vectors(p);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            mags[i * n + j] = m[i][j];
            for (size_t k = 0; k < 3; ++k) units[(i * n + j) * 3 + k] = u[i][j][k];
        }
}

// io::write_array (include/io.hpp:13-26), the writer of simulation::save_results
int ref_write_array(const char *fname, const double *arr, size_t len)
{
    return io::write_array(std::string(fname), arr, len);
}

double ref_field_square_fourier(double t, double h, double f, size_t n)
{
    return field::square_fourier(t, h, f, n);
}

}  // extern "C"
