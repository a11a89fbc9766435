// adjoint.hip -- host side of the pathwise sensitivities (mp_adjoint_* in
// include/magpy_amd.h): the SI -> reduced parameters of every particle and
// their derivatives with respect to theta = (K, alpha, H, median radius);
// the forward pass in blocks, each followed (when the gradient is recorded)
// by the transfers of its segments, computed in parallel and folded into
// one transfer per sampling interval; the reverse pass as a scan over the
// intervals; and the chain rule from per-particle reduced gradients to theta.
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

// lanes (segment x particle) per adjoint-map launch
int64_t env_batch_lanes()
{
    if (const char *env = std::getenv("MAGPY_AMD_ADJOINT_BATCH_LANES")) {
        long long v = std::atoll(env);
        if (v > 0) return (int64_t)v;
    }
    return 1 << 18;
}

struct AdjointBase {
    virtual ~AdjointBase() = default;
    virtual void begin(const double th[4], bool record) = 0;
    virtual int64_t forward_advance(int64_t n) = 0;
    virtual void forward_result(double *ens, double *fin, uint32_t *status) = 0;
    virtual void reverse_begin(const double *w, double wf) = 0;
    virtual int64_t reverse_advance(int64_t n) = 0;
    virtual void reverse_result(double grad[4]) = 0;
    virtual void particle_gradients(double *out) = 0;
    virtual void drain() = 0;
    double fwd_ms = 0, rev_ms = 0;
    int device = 0;
    int64_t ck_bytes = 0, scratch_bytes = 0;
};

// Pieces of the run: every segment [jC, min(jC + C, N)) is cut at the
// sample steps strictly inside it; a piece that starts at a sample step (or
// at step 0) starts a sampling interval.  Built once per handle.
struct PieceTable {
    std::vector<int64_t> bounds;                 // interval starts: 0 and the samples in (0, N)
    std::vector<int32_t> seg_first, seg_count;   // per segment
    std::vector<int64_t> start;                  // per piece: first step
    std::vector<int32_t> is_new, iv;             // per piece

    void build(const std::vector<int64_t> &sched, int64_t N, int64_t C)
    {
        bounds.assign(1, 0);
        for (int64_t k : sched)
            if (k > 0 && k < N && k != bounds.back()) bounds.push_back(k);
        const int64_t n_seg = (N + C - 1) / C;
        seg_first.resize((size_t)n_seg);
        seg_count.resize((size_t)n_seg);
        size_t bi = 0;   // next boundary to place
        for (int64_t j = 0; j < n_seg; ++j) {
            const int64_t a = j * C, b = std::min(N, a + C);
            seg_first[(size_t)j] = (int32_t)start.size();
            while (bi < bounds.size() && bounds[bi] < a) ++bi;
            auto add = [&](int64_t x) {
                // interval = last boundary <= x
                const size_t k = (size_t)(std::upper_bound(bounds.begin(), bounds.end(), x) -
                                          bounds.begin()) - 1;
                start.push_back(x);
                is_new.push_back(bounds[k] == x ? 1 : 0);
                iv.push_back((int32_t)k);
            };
            add(a);
            size_t bj = bi;
            while (bj < bounds.size() && bounds[bj] < b) {
                if (bounds[bj] > a) add(bounds[bj]);
                ++bj;
            }
            seg_count[(size_t)j] = (int32_t)(start.size() - seg_first[(size_t)j]);
        }
        require(start.size() < (size_t)INT32_MAX, "too many pieces");
    }
};

template <typename T>
struct AdjointRun final : AdjointBase {
    mp_adjoint_spec sp;
    std::vector<double> shape, m0;
    std::vector<int64_t> sched;
    PieceTable pt;
    int64_t P, N, C, S, n_ck, spl, segs_per_block, batch_segs, max_batch_pieces;
    int n_waves;
    // forward kernels on sg.s, the adjoint maps on tr.s: the two run
    // concurrently (the forward pass holds one wave per 64 particles; the
    // maps fill the rest of the machine)
    StreamGuard sg, tr;
    hipEvent_t ev_fwd[2] = {nullptr, nullptr}, ev_tr[2] = {nullptr, nullptr};
    bool tr_pending[2] = {false, false};
    int64_t ck_half = 0;   // checkpoint buffer of one block; two alternate
    DevBuf<T> ex, ey, ez, sig, mx, my, mz, ck, scratch, ftab, ftab_tr;
    DevBuf<uint32_t> key0, key1, status;
    DevBuf<double> partial, pieces, acc, intervals, inj, g;
    DevBuf<int64_t> d_sched, d_piece_start;
    DevBuf<int32_t> d_seg_first, d_seg_count, d_piece_new, d_piece_iv;
    // pass state
    std::vector<double> q, dq;
    bool begun = false, recording = false, fwd_finished = false, rev_begun = false, rev_done = false;
    int64_t fwd_done = 0;
    double lambda_z0 = 0.0;

    explicit AdjointRun(const mp_adjoint_spec &s) : sp(s)
    {
        device = s.device;
        P = sp.n_particles;
        N = sp.n_steps;
        C = sp.checkpoint_interval;
        S = sp.n_samples;
        n_ck = (N + C - 1) / C;
        spl = std::max<int64_t>(C, env_steps_per_launch() / C * C);
        segs_per_block = spl / C;
        shape.assign(sp.radius_shape, sp.radius_shape + P);
        m0.assign(sp.magnetisation_direction, sp.magnetisation_direction + 3 * P);
        sched.assign(sp.sample_steps, sp.sample_steps + S);
        sp.radius_shape = shape.data();
        sp.magnetisation_direction = m0.data();
        sp.sample_steps = sched.empty() ? nullptr : sched.data();
        sp.anisotropy_axis = nullptr;   // uploaded below, not kept
        sp.seed = nullptr;
        pt.build(sched, N, C);
        n_waves = (int)((P + mp::kWave - 1) / mp::kWave);
        // transfer batches: about 2^18 lanes (4,096 waves; 2^18 to 2^21
        // measure the same), scratch at most 8 GiB, the block's segments
        // split evenly
        const int64_t seg_len = std::min(C, N);
        const int64_t by_lanes = std::max<int64_t>(1, env_batch_lanes() / P);
        const int64_t by_bytes = std::max<int64_t>(
            1, ((int64_t)8 << 30) / (seg_len * mp::kScratchRows * P * (int64_t)sizeof(T)));
        const int64_t most = std::min({segs_per_block, by_lanes, by_bytes});
        const int64_t n_batches = (segs_per_block + most - 1) / most;
        batch_segs = (segs_per_block + n_batches - 1) / n_batches;
        max_batch_pieces = 1;
        for (int64_t b0 = 0; b0 < n_ck; b0 += segs_per_block) {
            const int64_t b1 = std::min(n_ck, b0 + segs_per_block);
            for (int64_t j0 = b0; j0 < b1; j0 += batch_segs) {
                const int64_t j1 = std::min(b1, j0 + batch_segs);
                const int64_t np = (j1 < n_ck ? pt.seg_first[(size_t)j1] : (int64_t)pt.start.size()) -
                                   pt.seg_first[(size_t)j0];
                max_batch_pieces = std::max(max_batch_pieces, np);
            }
        }
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
        g.alloc(4 * P);
        ck_half = (segs_per_block + 1) * 3 * P;
        ck.alloc((size_t)(2 * ck_half));
        ck_bytes = (int64_t)(ck.n * sizeof(T));
        for (int b = 0; b < 2; ++b) {
            MP_HIP(hipEventCreateWithFlags(&ev_fwd[b], hipEventDisableTiming));
            MP_HIP(hipEventCreateWithFlags(&ev_tr[b], hipEventDisableTiming));
        }
        partial.alloc((size_t)(S + 1) * n_waves);
        if (S) d_sched.upload(sched);
        ftab.alloc((size_t)spl + 5);  // the last 4-step group reads 4 past n1 + 1
        ftab_tr.alloc((size_t)spl + 2);
        d_seg_first.upload(pt.seg_first);
        d_seg_count.upload(pt.seg_count);
        d_piece_start.upload(pt.start);
        d_piece_new.upload(pt.is_new);
        d_piece_iv.upload(pt.iv);
        inj.alloc(pt.bounds.size());
    }

    ~AdjointRun() override
    {
        (void)hipStreamSynchronize(sg.s);
        (void)hipStreamSynchronize(tr.s);
        for (Timing &t : timings) {
            (void)hipEventDestroy(t.e0);
            (void)hipEventDestroy(t.e1);
        }
        for (hipEvent_t e : spare) (void)hipEventDestroy(e);
        for (int b = 0; b < 2; ++b) {
            if (ev_fwd[b]) (void)hipEventDestroy(ev_fwd[b]);
            if (ev_tr[b]) (void)hipEventDestroy(ev_tr[b]);
        }
    }

    // device buffers of the gradient recording, allocated on first use
    void alloc_recording()
    {
        if (scratch.p) return;
        const int64_t seg_len = std::min(C, N);
        scratch.alloc((size_t)seg_len * mp::kScratchRows * (size_t)(batch_segs * P));
        scratch_bytes = (int64_t)(scratch.n * sizeof(T));
        pieces.alloc((size_t)max_batch_pieces * mp::kTransfer * P);
        acc.alloc((size_t)mp::kTransfer * P);
        intervals.alloc(pt.bounds.size() * mp::kTransfer * (size_t)P);
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
        a.status = status.p;
        a.sched = d_sched.p;
        a.S = S;
        a.mz_partial = partial.p;
        a.n_waves = n_waves;
        return a;
    }

    void table(DevBuf<T> &tab, hipStream_t st, int64_t base, int64_t count)
    {
        require((size_t)count <= tab.n, "field table too small");
        const double fdt = sp.field_frequency * sp.time_step;
        hipLaunchKernelGGL(mp::path_field_table_kernel<T>, dim3((unsigned)((count + 255) / 256)),
                           dim3(256), 0, st, fdt, base, count, tab.p);
        MP_HIP(hipGetLastError());
    }

    // timing events of launches in flight: (stream's accumulator, start, end)
    struct Timing {
        double *acc;
        hipEvent_t e0, e1;
    };
    std::vector<Timing> timings;
    std::vector<hipEvent_t> spare;   // finished timing events, reused

    hipEvent_t timing_event()
    {
        hipEvent_t e = nullptr;
        if (!spare.empty()) {
            e = spare.back();
            spare.pop_back();
        } else {
            MP_HIP(hipEventCreate(&e));
        }
        return e;
    }

    template <typename F>
    void timed(double &acc, hipStream_t st, F &&f)
    {
        Timing t{&acc, timing_event(), timing_event()};
        timings.push_back(t);
        MP_HIP(hipEventRecord(t.e0, st));
        f();
        MP_HIP(hipEventRecord(t.e1, st));
    }

    // account the timed launches that have finished
    void collect()
    {
        size_t keep = 0;
        for (Timing &t : timings) {
            const hipError_t q = hipEventQuery(t.e1);
            if (q == hipErrorNotReady) {
                timings[keep++] = t;
                continue;
            }
            MP_HIP(q);
            float ms = 0.f;
            MP_HIP(hipEventElapsedTime(&ms, t.e0, t.e1));
            *t.acc += ms;
            spare.push_back(t.e0);
            spare.push_back(t.e1);
        }
        timings.resize(keep);
    }

    // wait for both streams and account every timed launch
    void drain() override
    {
        MP_HIP(hipStreamSynchronize(sg.s));
        MP_HIP(hipStreamSynchronize(tr.s));
        collect();
    }

    void begin(const double th[4], bool record) override
    {
        reduce_params(sp, th, q, dq);
        MP_HIP(hipSetDevice(device));
        // the previous pass's adjoint maps may still read sig and the state
        drain();
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
        // cleared on the forward stream, ahead of its kernels
        MP_HIP(hipMemsetAsync(status.p, 0, P * sizeof(uint32_t), sg.s));
        MP_HIP(hipMemsetAsync(partial.p, 0, partial.n * sizeof(double), sg.s));
        tr_pending[0] = tr_pending[1] = false;
        recording = record;
        if (recording) alloc_recording();
        begun = true;
        fwd_finished = false;
        rev_begun = false;
        rev_done = false;
        fwd_done = 0;
    }

    // transfers of segments [j0, j1) (all inside the current forward block,
    // whose first segment is ck_base), folded into the intervals
    void record_batch(int64_t j0, int64_t j1, int64_t ck_base, const T *ck_block)
    {
        const int64_t n_a = j0 * C, n_end = std::min(N, j1 * C);
        table(ftab_tr, tr.s, n_a, n_end - n_a + 2);
        const int64_t piece0 = pt.seg_first[(size_t)j0];
        const int64_t piece1 = j1 < n_ck ? pt.seg_first[(size_t)j1] : (int64_t)pt.start.size();
        mp::TransferArgs<T> t{};
        t.n = (int)P;
        t.ex = ex.p;
        t.ey = ey.p;
        t.ez = ez.p;
        t.sig = sig.p;
        t.key0 = key0.p;
        t.key1 = key1.p;
        t.tag = sp.noise_tag;
        t.dt = (T)q[0];
        t.alpha = (T)q[2];
        t.happ = (T)q[3];
        t.ftab = ftab_tr.p;
        t.tab0 = n_a;
        t.n_steps = N;
        t.ck_every = C;
        t.ck = ck_block + (size_t)(j0 - ck_base) * 3 * P;
        t.seg0 = j0;
        t.n_segs = j1 - j0;
        t.scratch = scratch.p;
        t.seg_first_piece = d_seg_first.p + j0;
        t.seg_n_pieces = d_seg_count.p + j0;
        t.piece_start = d_piece_start.p;
        t.piece_base = piece0;
        t.pieces = pieces.p;
        const int64_t lanes = (j1 - j0) * P;
        hipLaunchKernelGGL(mp::path_transfer_kernel<T>, dim3((unsigned)((lanes + 63) / 64)),
                           dim3(mp::kWave), 0, tr.s, t);
        MP_HIP(hipGetLastError());
        mp::ComposeArgs c{};
        c.n = (int)P;
        c.n_pieces = (int)(piece1 - piece0);
        c.pieces = pieces.p;
        c.piece_new = d_piece_new.p + piece0;
        c.piece_iv = d_piece_iv.p + piece0;
        c.acc = acc.p;
        c.intervals = intervals.p;
        c.close_last = j1 == n_ck ? 1 : 0;
        c.last_iv = (int)pt.bounds.size() - 1;
        hipLaunchKernelGGL(mp::path_compose_kernel, dim3((unsigned)n_waves), dim3(mp::kWave), 0,
                           tr.s, c);
        MP_HIP(hipGetLastError());
    }

    int64_t forward_advance(int64_t n) override
    {
        require(begun, "mp_adjoint_begin must come first");
        const int64_t target = std::min(N, fwd_done + std::max<int64_t>(n, 0));
        MP_HIP(hipSetDevice(device));
        mp::PathArgs<T> a = args();
        while (fwd_done < target) {
            // launches start on checkpoints (multiples of C, hence of 4)
            const int64_t n0 = fwd_done, n1 = std::min(N, n0 + spl);
            const int64_t j0 = n0 / C, j1 = (n1 + C - 1) / C;
            const int buf = (int)((n0 / spl) & 1);
            T *ck_block = ck.p + (size_t)buf * ck_half;
            // the maps of block k - 2 must be done with this checkpoint buffer
            if (tr_pending[buf]) MP_HIP(hipStreamWaitEvent(sg.s, ev_tr[buf], 0));
            timed(fwd_ms, sg.s, [&] {
                table(ftab, sg.s, n0, n1 - n0 + 5);
                a.tab0 = n0;
                a.n0 = n0;
                a.n1 = n1;
                a.ck = ck_block;
                a.ck_base = j0;
                a.s_next = std::lower_bound(sched.begin(), sched.end(), n0) - sched.begin();
                hipLaunchKernelGGL(mp::path_forward_kernel<T>, dim3((unsigned)n_waves),
                                   dim3(2 * mp::kWave), 0, sg.s, a);
                MP_HIP(hipGetLastError());
            });
            MP_HIP(hipEventRecord(ev_fwd[buf], sg.s));
            if (recording) {
                MP_HIP(hipStreamWaitEvent(tr.s, ev_fwd[buf], 0));
                timed(rev_ms, tr.s, [&] {
                    for (int64_t b = j0; b < j1; b += batch_segs)
                        record_batch(b, std::min(j1, b + batch_segs), j0, ck_block);
                });
                MP_HIP(hipEventRecord(ev_tr[buf], tr.s));
                tr_pending[buf] = true;
            }
            fwd_done = n1;
        }
        // the forward stream is done; the maps of the last block may still
        // run on the transfer stream, overlapping the next call's forward
        MP_HIP(hipStreamSynchronize(sg.s));
        collect();
        if (fwd_done == N) fwd_finished = true;
        return fwd_done;
    }

    void forward_result(double *ens, double *fin, uint32_t *st) override
    {
        if (!fwd_finished) forward_advance(N);
        drain();
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
        require(recording, "the forward pass did not record the gradient (record_gradient = 0)");
        if (!fwd_finished) forward_advance(N);
        require(S == 0 || w != nullptr, "grad_ens_mz is required when there are samples");
        // J = sum_s w_s Ms/P sum_p m_z,p(k_s) + wf/P sum_p m_z,p(N): the
        // weight of step k goes into lambda_z where the scan reaches step k
        std::vector<double> in(pt.bounds.size(), 0.0);
        lambda_z0 = wf / (double)P;
        for (int64_t s = 0; s < S; ++s) {
            const double c = w[s] * sp.magnetisation / (double)P;
            if (sched[s] == N) {
                lambda_z0 += c;
            } else {
                const size_t k = (size_t)(std::lower_bound(pt.bounds.begin(), pt.bounds.end(),
                                                           sched[s]) - pt.bounds.begin());
                in[k] += c;   // sched[s] is a boundary (or 0)
            }
        }
        MP_HIP(hipSetDevice(device));
        drain();
        MP_HIP(hipMemcpy(inj.p, in.data(), in.size() * sizeof(double), hipMemcpyHostToDevice));
        rev_begun = true;
        rev_done = false;
    }

    int64_t reverse_advance(int64_t) override
    {
        require(rev_begun, "mp_adjoint_reverse_begin must come first");
        if (!rev_done) {
            MP_HIP(hipSetDevice(device));
            timed(rev_ms, tr.s, [&] {
                mp::ScanArgs a{};
                a.n = (int)P;
                a.n_intervals = (int)pt.bounds.size();
                a.intervals = intervals.p;
                a.inj = inj.p;
                a.lambda_z0 = lambda_z0;
                a.g = g.p;
                hipLaunchKernelGGL(mp::path_scan_kernel, dim3((unsigned)n_waves), dim3(mp::kWave),
                                   0, tr.s, a);
                MP_HIP(hipGetLastError());
            });
            drain();
            rev_done = true;
        }
        return 0;
    }

    // per particle: dJ/dtheta [P][4] through that particle's trajectory
    void particle_gradients(double *out) override
    {
        require(rev_begun, "mp_adjoint_reverse_begin must come first");
        reverse_advance(0);
        std::vector<double> hg(g.n);
        MP_HIP(hipMemcpy(hg.data(), g.p, hg.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t p = 0; p < P; ++p) {
            const double *d = dq.data() + p * kQ * 4;
            for (int i = 0; i < 4; ++i) out[p * 4 + i] = 0.0;
            for (int j = 0; j < kQ; ++j) {
                const double gj = hg[(size_t)j * P + p];
                for (int i = 0; i < 4; ++i) out[p * 4 + i] += gj * d[j * 4 + i];
            }
        }
    }

    void reverse_result(double grad[4]) override
    {
        std::vector<double> per((size_t)P * 4);
        particle_gradients(per.data());
        for (int i = 0; i < 4; ++i) grad[i] = 0.0;
        for (int64_t p = 0; p < P; ++p)
            for (int i = 0; i < 4; ++i) grad[i] += per[(size_t)p * 4 + i];
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

int mp_adjoint_begin(mp_adjoint *a, const double theta[4], int record_gradient)
{
    return guarded([&] {
        require(a && theta, "handle and theta are required");
        a->impl->begin(theta, record_gradient != 0);
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

int mp_adjoint_particle_gradients(mp_adjoint *a, double *grad)
{
    return guarded([&] {
        require(a && grad, "handle and grad are required");
        a->impl->particle_gradients(grad);
    });
}

int mp_adjoint_info(mp_adjoint *a, double *forward_ms, double *reverse_ms,
                    int64_t *checkpoint_bytes, int64_t *scratch_bytes)
{
    return guarded([&] {
        require(a != nullptr, "handle is required");
        MP_HIP(hipSetDevice(a->impl->device));
        a->impl->drain();
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
