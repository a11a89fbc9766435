// simulation.hpp -- C++ entry of the MI355X sLLG ensemble integrator.
//
// simulation::ensemble_run is the batched counterpart of the reference's
// simulation::full_dynamics (include/simulation.hpp:74-94): for every member
// of the ensemble it has the semantics of one full_dynamics call with that
// member's arrays and seed, run as one set of device launches instead of one
// process per member (magpy/model.py:202-207).  Errors are thrown as
// simulation::Error with an MP_ERR_* code; the C ABI (magpy_amd.h) converts
// them to status codes.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/magpy_amd.h"

namespace simulation {

using EnsembleSpec = mp_ensemble_spec;
using EnsembleResult = mp_ensemble_out;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

// Reduced (dimensionless) parameters of one member, lib/simulation.cpp:501-549.
struct ReducedMember {
    double average_volume, average_anisotropy;
    double anisotropy_field;  // H_k = 2 K_avg / (mu0 Ms)
    double time_factor;       // gamma mu0 H_k / (1 + alpha^2)
    double dt, end_time;      // reduced
    double happ, fapp;        // reduced field amplitude and frequency
    std::vector<double> k_red, v_red, sr;
};

ReducedMember reduce_member(const double *radius, const double *anisotropy, size_t n,
                            double magnetisation, double damping, double temperature,
                            double time_step, double end_time, double field_amplitude,
                            double field_frequency);

// k[s] for s < max_samples and the number of steps the reference's loop takes.
int64_t sample_schedule(double dt, double end_time, int64_t max_samples, std::vector<int64_t> &k);

// throws Error(MP_ERR_ARG / MP_ERR_UNSUPPORTED) for requests it cannot run
void validate(const EnsembleSpec &spec);

void ensemble_run(const EnsembleSpec &spec, EnsembleResult &out);

}  // namespace simulation
