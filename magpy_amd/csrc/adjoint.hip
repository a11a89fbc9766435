// adjoint.hip -- host side of the pathwise sensitivities (mp_adjoint_* in
// include/magpy_amd.h): the SI -> reduced parameters of every particle and
// their derivatives with respect to theta = (K, alpha, H, median radius),
// the forward pass with device checkpoints, the reverse pass over segments,
// and the chain rule from the per-particle reduced gradients to theta.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "adjoint_kernels.hpp"
#include "host_util.hpp"

using namespace mp_host;

namespace {

constexpr int kQ = 4;  // reduced parameters per particle: dt, sig, alpha, happ

// Reduced parameters of particle p and dq_j/dtheta_i, as
// lib/simulation.cpp:501-549 reduces a one-particle member:
//   V = 4/3 pi r^3, H_k = 2K / (mu0 Ms), tau = gamma mu0 H_k / (1 + alpha^2),
//   dt = time_step tau, sr = sqrt(alpha kB T / (K V) / (1 + alpha^2)),
//   sig = sr sqrt(dt), happ = H / H_k.
// sig^2 = 2 alpha kB T time_step gamma / (V Ms (1 + alpha^2)^2) does not
// depend on K.
void reduce_params(const mp_adjoint_spec &sp, const double th[4], std::vector<double> &q,
                   std::vector<double> &dq)
{
#pragma clang fp contract(off)
    const double K = th[0], al = th[1], H = th[2], rm = th[3];
    require(K > 0 && al > 0 && rm > 0 && std::isfinite(H),
            "theta needs anisotropy > 0, damping > 0, radius > 0 and a finite field");
    const size_t P = (size_t)sp.n_particles;
    q.assign(P * kQ, 0.0);
    dq.assign(P * kQ * 4, 0.0);
    const double hk = 2 * K / kMU0 / sp.magnetisation;
    const double tf = kGYROMAG * kMU0 * hk / (1 + al * al);
    const double dt = sp.time_step * tf;
    const double happ = H / hk;
    const double dlog_dt_dal = -2 * al / (1 + al * al);
    for (size_t p = 0; p < P; ++p) {
        const double r = rm * sp.radius_shape[p];
        const double vol = 4.0 / 3.0 * kPiHost * r * r * r;
        const double sr = std::sqrt(al * kKB * sp.temperature / (K * vol) / (1 + al * al));
        const double sig = sr * std::sqrt(dt);
        double *qp = q.data() + p * kQ;
        double *d = dq.data() + p * kQ * 4;   // d[j * 4 + i] = dq_j / dtheta_i
        qp[0] = dt;
        qp[1] = sig;
        qp[2] = al;
        qp[3] = happ;
        d[0 * 4 + 0] = dt / K;
        d[0 * 4 + 1] = dt * dlog_dt_dal;
        d[1 * 4 + 1] = sig * (0.5 / al + dlog_dt_dal);
        d[1 * 4 + 3] = -1.5 * sig / rm;
        d[2 * 4 + 1] = 1.0;
        d[3 * 4 + 0] = -happ / K;
        d[3 * 4 + 2] = 1.0 / hk;
    }
}

void validate(const mp_adjoint_spec &sp)
{
    require(sp.n_particles > 0 && sp.n_particles <= (1 << 30), "n_particles must be in [1, 2^30]");
    require(sp.radius_shape && sp.anisotropy_axis && sp.magnetisation_direction && sp.seed,
            "radius_shape, anisotropy_axis, magnetisation_direction and seed are required");
    require(sp.magnetisation > 0 && sp.temperature >= 0 && sp.time_step > 0 &&
                std::isfinite(sp.field_frequency),
            "magnetisation > 0, temperature >= 0, time_step > 0 and a finite frequency");
    require(sp.n_steps > 0, "n_steps must be positive");
    require(sp.checkpoint_interval > 0 && sp.checkpoint_interval % 4 == 0,
            "checkpoint_interval must be a positive multiple of 4");
    require(sp.n_samples >= 0 && (sp.n_samples == 0 || sp.sample_steps),
            "sample_steps is required when n_samples > 0");
    for (int64_t s = 0; s < sp.n_samples; ++s) {
        require(sp.sample_steps[s] >= 0 && sp.sample_steps[s] <= sp.n_steps,
                "sample_steps must lie in [0, n_steps]");
        require(s == 0 || sp.sample_steps[s] >= sp.sample_steps[s - 1],
                "sample_steps must be ascending");
    }
    for (int64_t p = 0; p < sp.n_particles; ++p)
        require(sp.radius_shape[p] > 0, "radius_shape must be positive");
    require(sp.precision == MP_PRECISION_FP32 || sp.precision == MP_PRECISION_FP64,
            "precision must be MP_PRECISION_FP32 or MP_PRECISION_FP64");
}

int64_t env_steps_per_launch()
{
    if (const char *env = std::getenv("MAGPY_AMD_ADJOINT_STEPS_PER_LAUNCH")) {
        long long v = std::atoll(env);
        if (v > 0) return (int64_t)v;
    }
    return 1 << 16;
}

struct AdjointBase {
    virtual ~AdjointBase() = default;
    virtual void begin(const double th[4]) = 0;
    virtual int64_t forward_advance(int64_t n) = 0;
    virtual void forward_result(double *ens, double *fin, uint32_t *status) = 0;
    virtual void reverse_begin(const double *w, double wf) = 0;
    virtual int64_t reverse_advance(int64_t n) = 0;
    virtual void reverse_result(double grad[4]) = 0;
    double fwd_ms = 0, rev_ms = 0;
    int64_t ck_bytes = 0, scratch_bytes = 0;
};

template <typename T>
struct AdjointRun final : AdjointBase {
    mp_adjoint_spec sp;
    std::vector<double> shape, m0;
    std::vector<int64_t> sched;
    int device;
    int64_t P, N, C, S, n_ck, spl;
    int n_waves;
    StreamGuard sg;
    DevBuf<T> ex, ey, ez, sig, mx, my, mz, ck, scratch, ftab, lam;
    DevBuf<uint32_t> key0, key1, status;
    DevBuf<double> partial, c_sample, g;
    DevBuf<int64_t> d_sched;
    // pass state
    double theta[4] = {0, 0, 0, 0};
    std::vector<double> q, dq;
    bool begun = false, fwd_finished = false, rev_begun = false;
    int64_t fwd_done = 0, seg_next = -1;

    explicit AdjointRun(const mp_adjoint_spec &s) : sp(s), device(s.device)
    {
        P = sp.n_particles;
        N = sp.n_steps;
        C = sp.checkpoint_interval;
        S = sp.n_samples;
        n_ck = (N + C - 1) / C;
        spl = std::max<int64_t>(C, env_steps_per_launch() / C * C);
        shape.assign(sp.radius_shape, sp.radius_shape + P);
        m0.assign(sp.magnetisation_direction, sp.magnetisation_direction + 3 * P);
        sched.assign(sp.sample_steps, sp.sample_steps + S);
        sp.radius_shape = shape.data();
        sp.magnetisation_direction = m0.data();
        sp.sample_steps = sched.empty() ? nullptr : sched.data();
        sp.anisotropy_axis = nullptr;   // uploaded below, not kept
        sp.seed = nullptr;
        n_waves = (int)((P + mp::kWave - 1) / mp::kWave);
        std::vector<T> hx(P), hy(P), hz(P);
        std::vector<uint32_t> k0(P), k1(P);
        for (int64_t p = 0; p < P; ++p) {
            hx[p] = (T)s.anisotropy_axis[3 * p];
            hy[p] = (T)s.anisotropy_axis[3 * p + 1];
            hz[p] = (T)s.anisotropy_axis[3 * p + 2];
            const uint64_t sd = (uint64_t)s.seed[p];
            k0[p] = (uint32_t)sd;
            k1[p] = (uint32_t)(sd >> 32);
        }
        MP_HIP(hipSetDevice(device));
        ex.upload(hx);
        ey.upload(hy);
        ez.upload(hz);
        key0.upload(k0);
        key1.upload(k1);
        sig.alloc(P);
        mx.alloc(P);
        my.alloc(P);
        mz.alloc(P);
        status.alloc(P);
        lam.alloc(3 * P);
        g.alloc(4 * P);
        ck.alloc((size_t)n_ck * 3 * P);
        scratch.alloc((size_t)std::min(C, N) * 3 * P);
        ck_bytes = (int64_t)(ck.n * sizeof(T));
        scratch_bytes = (int64_t)(scratch.n * sizeof(T));
        partial.alloc((size_t)(S + 1) * n_waves);
        c_sample.alloc(std::max<int64_t>(S, 1));
        if (S) d_sched.upload(sched);
        const int64_t tab = std::max(spl, std::min(C, N) * std::max<int64_t>(1, spl / C)) + 2;
        ftab.alloc((size_t)tab);
    }

    mp::PathArgs<T> args() const
    {
        mp::PathArgs<T> a{};
        a.n = (int)P;
        a.ex = ex.p;
        a.ey = ey.p;
        a.ez = ez.p;
        a.sig = sig.p;
        a.key0 = key0.p;
        a.key1 = key1.p;
        a.tag = sp.noise_tag;
        a.dt = (T)q[0];
        a.alpha = (T)q[2];
        a.happ = (T)q[3];
        a.mx = mx.p;
        a.my = my.p;
        a.mz = mz.p;
        a.ftab = ftab.p;
        a.n_steps = N;
        a.ck_every = C;
        a.ck = ck.p;
        a.scratch = scratch.p;
        a.status = status.p;
        a.sched = d_sched.p;
        a.S = S;
        a.mz_partial = partial.p;
        a.n_waves = n_waves;
        a.c_sample = c_sample.p;
        a.lx = lam.p;
        a.ly = lam.p + P;
        a.lz = lam.p + 2 * P;
        a.g = g.p;
        return a;
    }

    void table(int64_t base, int64_t count)
    {
        require((size_t)count <= ftab.n, "field table too small");
        const double fdt = sp.field_frequency * sp.time_step;
        hipLaunchKernelGGL(mp::path_field_table_kernel<T>, dim3((unsigned)((count + 255) / 256)),
                           dim3(256), 0, sg.s, fdt, base, count, ftab.p);
        MP_HIP(hipGetLastError());
    }

    template <typename F>
    double timed(F &&f)
    {
        MP_HIP(hipEventRecord(sg.e0, sg.s));
        f();
        MP_HIP(hipEventRecord(sg.e1, sg.s));
        MP_HIP(hipStreamSynchronize(sg.s));
        float ms = 0.f;
        MP_HIP(hipEventElapsedTime(&ms, sg.e0, sg.e1));
        return ms;
    }

    void begin(const double th[4]) override
    {
        reduce_params(sp, th, q, dq);
        std::copy(th, th + 4, theta);
        MP_HIP(hipSetDevice(device));
        std::vector<T> hs(P), hx(P), hy(P), hz(P);
        for (int64_t p = 0; p < P; ++p) {
            hs[p] = (T)q[p * kQ + 1];
            hx[p] = (T)m0[3 * p];
            hy[p] = (T)m0[3 * p + 1];
            hz[p] = (T)m0[3 * p + 2];
        }
        MP_HIP(hipMemcpy(sig.p, hs.data(), P * sizeof(T), hipMemcpyHostToDevice));
        MP_HIP(hipMemcpy(mx.p, hx.data(), P * sizeof(T), hipMemcpyHostToDevice));
        MP_HIP(hipMemcpy(my.p, hy.data(), P * sizeof(T), hipMemcpyHostToDevice));
        MP_HIP(hipMemcpy(mz.p, hz.data(), P * sizeof(T), hipMemcpyHostToDevice));
        MP_HIP(hipMemset(status.p, 0, P * sizeof(uint32_t)));
        MP_HIP(hipMemset(partial.p, 0, partial.n * sizeof(double)));
        begun = true;
        fwd_finished = false;
        rev_begun = false;
        fwd_done = 0;
    }

    int64_t forward_advance(int64_t n) override
    {
        require(begun, "mp_adjoint_begin must come first");
        const int64_t target = std::min(N, fwd_done + std::max<int64_t>(n, 0));
        MP_HIP(hipSetDevice(device));
        fwd_ms += timed([&] {
            mp::PathArgs<T> a = args();
            while (fwd_done < target) {
                // launches start on checkpoints (multiples of C, hence of 4)
                const int64_t n0 = fwd_done, n1 = std::min(N, n0 + spl);
                table(n0, n1 - n0 + 2);
                a.tab0 = n0;
                a.n0 = n0;
                a.n1 = n1;
                a.s_next = std::lower_bound(sched.begin(), sched.end(), n0) - sched.begin();
                hipLaunchKernelGGL(mp::path_forward_kernel<T>, dim3((unsigned)n_waves),
                                   dim3(mp::kWave), 0, sg.s, a);
                MP_HIP(hipGetLastError());
                fwd_done = n1;
            }
        });
        if (fwd_done == N) fwd_finished = true;
        return fwd_done;
    }

    void forward_result(double *ens, double *fin, uint32_t *st) override
    {
        if (!fwd_finished) forward_advance(N);
        std::vector<double> part(partial.n);
        MP_HIP(hipMemcpy(part.data(), partial.p, part.size() * sizeof(double),
                         hipMemcpyDeviceToHost));
        for (int64_t s = 0; s <= S; ++s) {
            double sum = 0.0;
            for (int w = 0; w < n_waves; ++w) sum += part[(size_t)s * n_waves + w];
            if (s < S && ens) ens[s] = sp.magnetisation * sum / (double)P;
            if (s == S && fin) *fin = sum / (double)P;
        }
        if (st) {
            std::vector<uint32_t> h(P);
            MP_HIP(hipMemcpy(h.data(), status.p, P * sizeof(uint32_t), hipMemcpyDeviceToHost));
            uint32_t o = 0;
            for (uint32_t v : h) o |= v;
            *st = o;
        }
    }

    void reverse_begin(const double *w, double wf) override
    {
        require(begun, "mp_adjoint_begin must come first");
        if (!fwd_finished) forward_advance(N);
        require(S == 0 || w != nullptr, "grad_ens_mz is required when there are samples");
        MP_HIP(hipSetDevice(device));
        // J = sum_s w_s Ms/P sum_p m_z,p(k_s) + wf/P sum_p m_z,p(N)
        std::vector<double> c((size_t)std::max<int64_t>(S, 1), 0.0);
        double lz0 = wf / (double)P;
        for (int64_t s = 0; s < S; ++s) {
            c[s] = w[s] * sp.magnetisation / (double)P;
            if (sched[s] == N) lz0 += c[s];
        }
        MP_HIP(hipMemcpy(c_sample.p, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
        std::vector<T> l(3 * P, T(0));
        for (int64_t p = 0; p < P; ++p) l[2 * P + p] = (T)lz0;
        MP_HIP(hipMemcpy(lam.p, l.data(), l.size() * sizeof(T), hipMemcpyHostToDevice));
        MP_HIP(hipMemset(g.p, 0, g.n * sizeof(double)));
        seg_next = n_ck - 1;
        rev_begun = true;
    }

    int64_t steps_left() const { return seg_next < 0 ? 0 : std::min(N, (seg_next + 1) * C); }

    int64_t reverse_advance(int64_t n) override
    {
        require(rev_begun, "mp_adjoint_reverse_begin must come first");
        const int64_t stop = std::max<int64_t>(0, steps_left() - std::max<int64_t>(n, 0));
        const int64_t segs = std::max<int64_t>(1, spl / C);
        MP_HIP(hipSetDevice(device));
        rev_ms += timed([&] {
            mp::PathArgs<T> a = args();
            while (seg_next >= 0 && steps_left() > stop) {
                const int64_t hi = seg_next, lo = std::max<int64_t>(0, hi - segs + 1);
                const int64_t n_lo = lo * C, n_end = std::min(N, (hi + 1) * C);
                table(n_lo, n_end - n_lo + 2);
                a.tab0 = n_lo;
                a.seg_lo = lo;
                a.seg_hi = hi;
                // last sample strictly before n_end (samples at n_end are in lambda)
                a.s_last = (int64_t)(std::lower_bound(sched.begin(), sched.end(), n_end) -
                                     sched.begin()) - 1;
                hipLaunchKernelGGL(mp::path_reverse_kernel<T>, dim3((unsigned)n_waves),
                                   dim3(mp::kWave), 0, sg.s, a);
                MP_HIP(hipGetLastError());
                seg_next = lo - 1;
            }
        });
        return steps_left();
    }

    void reverse_result(double grad[4]) override
    {
        require(rev_begun, "mp_adjoint_reverse_begin must come first");
        if (seg_next >= 0) reverse_advance(N);
        std::vector<double> hg(g.n);
        MP_HIP(hipMemcpy(hg.data(), g.p, hg.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < 4; ++i) grad[i] = 0.0;
        for (int64_t p = 0; p < P; ++p) {
            const double *d = dq.data() + p * kQ * 4;
            for (int j = 0; j < kQ; ++j) {
                const double gj = hg[(size_t)j * P + p];
                for (int i = 0; i < 4; ++i) grad[i] += gj * d[j * 4 + i];
            }
        }
    }
};

}  // namespace

struct mp_adjoint {
    std::unique_ptr<AdjointBase> impl;
};

extern "C" {

int mp_adjoint_create(const mp_adjoint_spec *spec, mp_adjoint **out)
{
    return guarded([&] {
        require(spec && out, "spec and out are required");
        *out = nullptr;
        validate(*spec);
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
            throw simulation::Error(MP_ERR_NODEV, "no HIP device visible");
        require(spec->device >= 0 && spec->device < count, "device index out of range");
        std::unique_ptr<mp_adjoint> a(new mp_adjoint);
        if (spec->precision == MP_PRECISION_FP32)
            a->impl.reset(new AdjointRun<float>(*spec));
        else
            a->impl.reset(new AdjointRun<double>(*spec));
        *out = a.release();
    });
}

void mp_adjoint_destroy(mp_adjoint *a) { delete a; }

int mp_adjoint_begin(mp_adjoint *a, const double theta[4])
{
    return guarded([&] {
        require(a && theta, "handle and theta are required");
        a->impl->begin(theta);
    });
}

int mp_adjoint_forward_advance(mp_adjoint *a, int64_t max_steps, int64_t *steps_done)
{
    return guarded([&] {
        require(a != nullptr, "handle is required");
        const int64_t d = a->impl->forward_advance(max_steps);
        if (steps_done) *steps_done = d;
    });
}

int mp_adjoint_forward_result(mp_adjoint *a, double *ens_mz, double *final_mz, uint32_t *status)
{
    return guarded([&] {
        require(a != nullptr, "handle is required");
        a->impl->forward_result(ens_mz, final_mz, status);
    });
}

int mp_adjoint_reverse_begin(mp_adjoint *a, const double *grad_ens_mz, double grad_final_mz)
{
    return guarded([&] {
        require(a != nullptr, "handle is required");
        a->impl->reverse_begin(grad_ens_mz, grad_final_mz);
    });
}

int mp_adjoint_reverse_advance(mp_adjoint *a, int64_t max_steps, int64_t *steps_left)
{
    return guarded([&] {
        require(a != nullptr, "handle is required");
        const int64_t l = a->impl->reverse_advance(max_steps);
        if (steps_left) *steps_left = l;
    });
}

int mp_adjoint_reverse_result(mp_adjoint *a, double grad_theta[4])
{
    return guarded([&] {
        require(a && grad_theta, "handle and grad_theta are required");
        a->impl->reverse_result(grad_theta);
    });
}

int mp_adjoint_info(mp_adjoint *a, double *forward_ms, double *reverse_ms,
                    int64_t *checkpoint_bytes, int64_t *scratch_bytes)
{
    return guarded([&] {
        require(a != nullptr, "handle is required");
        if (forward_ms) *forward_ms = a->impl->fwd_ms;
        if (reverse_ms) *reverse_ms = a->impl->rev_ms;
        if (checkpoint_bytes) *checkpoint_bytes = a->impl->ck_bytes;
        if (scratch_bytes) *scratch_bytes = a->impl->scratch_bytes;
    });
}

int mp_adjoint_reduced(const mp_adjoint_spec *spec, const double theta[4], double *q, double *dq)
{
    return guarded([&] {
        require(spec && theta && q && dq, "spec, theta, q and dq are required");
        require(spec->n_particles > 0 && spec->radius_shape, "particles are required");
        require(spec->magnetisation > 0 && spec->time_step > 0, "magnetisation and time_step > 0");
        std::vector<double> hq, hdq;
        reduce_params(*spec, theta, hq, hdq);
        std::copy(hq.begin(), hq.end(), q);
        std::copy(hdq.begin(), hdq.end(), dq);
    });
}

}  // extern "C"
