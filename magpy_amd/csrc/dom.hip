// dom.hip -- host side of the discrete orientation model (mp_dom_run in
// include/magpy_amd.h): validates the ensemble, reduces the field amplitude
// by each particle's anisotropy field as core.simulate_dom does
// (magpy/core.pyx:205-249), runs dom_kernel (dom_kernels.hpp) over all
// particles in one launch and returns the samples particle-major.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "dom_kernels.hpp"
#include "host_util.hpp"

using namespace mp_host;

namespace {

double dom_field_host(int shape, double h, double f, int64_t n, double t)
{
#pragma clang fp contract(off)
    switch (shape) {
    case MP_FIELD_SINE: return h * std::sin(2 * kPiHost * f * t);
    case MP_FIELD_SQUARE: return h * ((long long)(t * f * 2) % 2 ? -1.0 : 1.0);
    case MP_FIELD_SQUARE_FOURIER: {
        double s = 0.0;
        for (int64_t k = 1; k < n + 1; ++k) {
            const double odd = (double)(2 * k - 1);
            s += std::sin(2 * kPiHost * odd * f * t) / odd;
        }
        return s * (4 / kPiHost * h);
    }
    default: return h;
    }
}

void dom_run(const mp_dom_spec &sp, double *time, double *field, double *mz, uint32_t *status,
             int64_t *steps, double *kernel_ms)
{
#pragma clang fp contract(off)
    const int64_t P = sp.n_particles, S = sp.max_samples;
    require(P > 0 && P <= (1 << 30), "n_particles must be in [1, 2^30]");
    require(S >= 2, "max_samples must be >= 2");
    require(sp.volume && sp.anisotropy && sp.temperature && sp.magnetisation && sp.damping &&
                sp.initial_probabilities,
            "volume, anisotropy, temperature, magnetisation, damping and "
            "initial_probabilities are required");
    require(sp.field_shape >= MP_FIELD_SINE && sp.field_shape <= MP_FIELD_SQUARE_FOURIER,
            "unknown field shape");
    require(sp.field_shape != MP_FIELD_SQUARE_FOURIER || sp.field_n_components >= 1,
            "field_n_components must be >= 1");
    require(sp.time_step > 0 && sp.end_time > 0, "time_step and end_time must be positive");
    for (int64_t p = 0; p < P; ++p)
        require(sp.volume[p] > 0 && sp.anisotropy[p] > 0 && sp.temperature[p] > 0 &&
                    sp.magnetisation[p] > 0 && sp.damping[p] > 0,
                "particle parameters must be positive");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        throw simulation::Error(MP_ERR_NODEV, "no HIP device visible");
    require(sp.device >= 0 && sp.device < count, "device index out of range");
    MP_HIP(hipSetDevice(sp.device));

    std::vector<double> happ(P), up(P), down(P), hk(P);
    for (int64_t p = 0; p < P; ++p) {
        hk[p] = 2.0 * sp.anisotropy[p] / kMU0 / sp.magnetisation[p];
        happ[p] = sp.field_amplitude / hk[p];
        up[p] = sp.initial_probabilities[2 * p];
        down[p] = sp.initial_probabilities[2 * p + 1];
    }
    auto vec = [&](const double *x) { return std::vector<double>(x, x + P); };
    DevBuf<double> dk, dv, dt, dms, dal, dh, dup, ddown, dmz;
    dk.upload(vec(sp.anisotropy));
    dv.upload(vec(sp.volume));
    dt.upload(vec(sp.temperature));
    dms.upload(vec(sp.magnetisation));
    dal.upload(vec(sp.damping));
    dh.upload(happ);
    dup.upload(up);
    ddown.upload(down);
    dmz.alloc((size_t)S * P);
    DevBuf<uint32_t> dst(P);
    DevBuf<int64_t> dsteps(P);
    mp::DomArgs a{};
    a.n = (int)P;
    a.k = dk.p;
    a.v = dv.p;
    a.temp = dt.p;
    a.ms = dms.p;
    a.alpha = dal.p;
    a.happ = dh.p;
    a.p_up = dup.p;
    a.p_down = ddown.p;
    a.shape = sp.field_shape;
    a.freq = sp.field_frequency;
    a.n_components = sp.field_n_components;
    a.time_step = sp.time_step;
    a.end_time = sp.end_time;
    a.S = S;
    a.max_steps = sp.max_steps > 0 ? sp.max_steps : (int64_t)1 << 32;
    a.mz = dmz.p;
    a.status = dst.p;
    a.steps = dsteps.p;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MP_HIP(hipEventCreate(&e0));
    MP_HIP(hipEventCreate(&e1));
    MP_HIP(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(mp::dom_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64), 0, nullptr, a);
    MP_HIP(hipGetLastError());
    MP_HIP(hipEventRecord(e1, nullptr));
    MP_HIP(hipDeviceSynchronize());
    float ms = 0.f;
    MP_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (kernel_ms) *kernel_ms = ms;
    std::vector<double> h((size_t)S * P);
    MP_HIP(hipMemcpy(h.data(), dmz.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    if (mz)
        for (int64_t p = 0; p < P; ++p)
            for (int64_t s = 0; s < S; ++s) mz[p * S + s] = h[(size_t)s * P + p];
    if (status) MP_HIP(hipMemcpy(status, dst.p, P * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (steps) MP_HIP(hipMemcpy(steps, dsteps.p, P * sizeof(int64_t), hipMemcpyDeviceToHost));
    // time and field of particle 0 as the reference returns them: the
    // reduced field times that particle's anisotropy field
    const double ts = sp.end_time / (double)(S - 1);
    for (int64_t s = 0; s < S; ++s) {
        const double t = s == 0 ? 0.0 : (double)s * ts;
        if (time) time[s] = t;
        if (field)
            field[s] = dom_field_host(sp.field_shape, happ[0], sp.field_frequency,
                                      sp.field_n_components, t) * hk[0];
    }
}

}  // namespace

extern "C" int mp_dom_run(const mp_dom_spec *spec, double *time, double *field, double *mz,
                          uint32_t *status, int64_t *steps, double *kernel_ms)
{
    return guarded([&] {
        require(spec != nullptr, "spec is required");
        dom_run(*spec, time, field, mz, status, steps, kernel_ms);
    });
}
