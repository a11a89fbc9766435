// ensemble_run.hip -- host side of the MI355X sLLG integrator and its C ABI.
//
// simulation::ensemble_run (simulation.hpp) validates the request, reduces
// every member to dimensionless units exactly as simulation::full_dynamics
// does (lib/simulation.cpp:501-583), derives the reference's sample schedule,
// uploads SoA parameter arrays, and drives the step kernels of
// sllg_kernels.hpp in blocks of steps on one HIP stream.  A session keeps that
// state resident between calls.  The extern "C" functions at the end are the
// boundary declared in include/magpy_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "host_util.hpp"
#include "simulation.hpp"
#include "replica_kernels.hpp"
#include "sllg_kernels.hpp"

namespace mp_host {
std::string &last_error()
{
    thread_local std::string text;
    return text;
}
}  // namespace mp_host

namespace {

using namespace mp_host;

// The applied field of lib/simulation.cpp:552-573 at reduced time t.
double applied_host(int shape, double happ, double fapp, double t)
{
    switch (shape) {
    case MP_FIELD_SINE: return happ * std::sin(2 * kPiHost * fapp * t);
    case MP_FIELD_SQUARE: return happ * ((int)(t * fapp * 2) % 2 ? -1 : 1);
    default: return happ;
    }
}

// Steps per kernel launch: long enough to amortise the launch, short enough
// to keep each launch well under a second.
int64_t steps_per_launch(int scheme, int precision)
{
    if (const char *env = std::getenv("MAGPY_AMD_STEPS_PER_LAUNCH")) {
        long long v = std::atoll(env);
        if (v > 0) return (int64_t)v;
    }
    if (scheme == MP_SCHEME_IMPLICIT_MIDPOINT) return 1 << 12;
    return precision == MP_PRECISION_FP32 ? (1 << 17) : (1 << 16);
}

template <typename T>
std::vector<T> cast_vec(const std::vector<double> &v)
{
    return std::vector<T>(v.begin(), v.end());
}

// Which Heun kernel runs n_waves integrating waves: 0 the one-wave kernel,
// 1 the two-wave kernel with the normals formed by the producing wave, 2 (FP32)
// the two-wave kernel with Box-Muller in the integrating wave.  FP32: 2 while
// the integrating waves are more than one per CU and at most one per SIMD, 1
// at one per CU or fewer, else 0 (32,768-step launches: 65,536 particles
// 3.63e11 / 3.02e11 / 2.72e11 for 2 / 1 / 0; 16,384: 0.94e11 / 1.21e11 /
// 0.68e11; 131,072: 2.99e11 / 2.82e11 / 3.66e11).  FP64: 0, the one-wave
// kernel is faster at every size measured (65,536 particles: 5.7e10 against
// 4.3e10; 16,384: 1.50e10 against 1.44e10; 1,000: 1.37e9 against 1.20e9).
// MAGPY_AMD_HEUN_TWO_WAVE=0/1/2 forces the choice (2 means 1 in FP64).
int two_wave_mode(int n_waves, bool fp32)
{
    if (const char *env = std::getenv("MAGPY_AMD_HEUN_TWO_WAVE")) {
        const int v = std::atoi(env);
        return v <= 0 ? 0 : (v >= 2 && fp32) ? 2 : 1;
    }
    if (!fp32) return 0;
    int dev = 0, cus = 0;
    MP_HIP(hipGetDevice(&dev));
    MP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (n_waves > 4 * cus) return 0;
    return n_waves > cus ? 2 : 1;
}

// grid: blocks of kBlock lanes (n_waves = 4 grid integrating waves)
template <typename T, int kShape, bool kArray, bool kRenorm>
void launch_heun(const mp::IndepArgs<T> &a, int grid, hipStream_t s, const float *htab,
                 int two_wave)
{
    if constexpr (!kArray) {
        if (two_wave != 0) {
            // one (integrating, producing) wave pair per workgroup; two or
            // four pairs per workgroup measure the same or slower
            const dim3 g2((unsigned)(grid * (mp::kBlock / mp::kWave))), b2(2 * mp::kWave);
            if constexpr (sizeof(T) == 4) {
                auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, g2, b2, 0, s, a, htab); };
                if (htab) {
                    if (two_wave == 2)
                        go(mp::heun_indep_pc_kernel<T, mp::kConstant, true, kRenorm, true>);
                    else
                        go(mp::heun_indep_pc_kernel<T, mp::kConstant, true, kRenorm, false>);
                    return;
                }
                if (two_wave == 2) go(mp::heun_indep_pc_kernel<T, kShape, false, kRenorm, true>);
                else go(mp::heun_indep_pc_kernel<T, kShape, false, kRenorm, false>);
                return;
            }
            hipLaunchKernelGGL((mp::heun_indep_pc_kernel<T, kShape, false, kRenorm, false>), g2,
                               b2, 0, s, a, htab);
            return;
        }
    }
    if constexpr (sizeof(T) == 4) {
        if (htab) {  // the table replaces per-lane field evaluation: one variant per noise kind
            hipLaunchKernelGGL((mp::heun_indep_kernel<T, mp::kConstant, kArray, true, kRenorm>),
                               dim3(grid), dim3(mp::kBlock), 0, s, a, htab);
            return;
        }
    }
    hipLaunchKernelGGL((mp::heun_indep_kernel<T, kShape, kArray, false, kRenorm>), dim3(grid),
                       dim3(mp::kBlock), 0, s, a, htab);
}

template <typename T, int kShape, bool kArray>
void launch_step(const mp::IndepArgs<T> &a, int scheme, double eps, int max_iter,
                 unsigned long long *it, unsigned long long *fl, int grid, hipStream_t s,
                 const float *htab, int two_wave)
{
    if constexpr (sizeof(T) == 8) {
        if (scheme == MP_SCHEME_IMPLICIT_MIDPOINT) {
            mp::ImplicitArgs ia{a, eps, max_iter, it, fl};
            hipLaunchKernelGGL((mp::implicit_indep_kernel<kShape, kArray>), dim3(grid),
                               dim3(mp::kBlock), 0, s, ia);
            return;
        }
    }
    if (a.renorm)
        launch_heun<T, kShape, kArray, true>(a, grid, s, htab, two_wave);
    else
        launch_heun<T, kShape, kArray, false>(a, grid, s, htab, two_wave);
}

template <typename T, bool kArray>
void launch_shape(int shape_kind, const mp::IndepArgs<T> &a, int scheme, double eps,
                  int max_iter, unsigned long long *it, unsigned long long *fl, int grid,
                  hipStream_t s, const float *htab, int tw)
{
    switch (shape_kind) {
    case mp::kSine: launch_step<T, mp::kSine, kArray>(a, scheme, eps, max_iter, it, fl, grid, s, htab, tw); break;
    case mp::kSquare: launch_step<T, mp::kSquare, kArray>(a, scheme, eps, max_iter, it, fl, grid, s, htab, tw); break;
    case mp::kConstant: launch_step<T, mp::kConstant, kArray>(a, scheme, eps, max_iter, it, fl, grid, s, htab, tw); break;
    default: launch_step<T, mp::kMixed, kArray>(a, scheme, eps, max_iter, it, fl, grid, s, htab, tw); break;
    }
}

struct Plan {
    // particles ordered by schedule group; perm[i] = original particle index
    std::vector<int64_t> perm;
    std::vector<int64_t> group_begin;        // [G+1] ranges in perm order
    std::vector<std::vector<int64_t>> sched; // per group
    std::vector<int64_t> group_steps;        // per group
    std::vector<int64_t> member_of;          // per original particle
    std::vector<simulation::ReducedMember> red;
};

// Host copies of what the outputs need after the caller's spec is gone.
struct HostCopy {
    int64_t M = 0, P = 0, S = 0, start = 0;   // m0 is the state at step `start`
    std::vector<double> m0, ms;
    std::vector<int32_t> shape;
};

// Results of one group copied back to the host, per slot.
struct GroupOut {
    std::vector<int64_t> particle_of; // slot -> particle (-1 = padding)
    std::vector<double> state;        // [slots][3]
    std::vector<uint32_t> status;
    std::vector<double> samples;      // [S][3][slots] (empty when not recorded)
    std::vector<double> ens;          // [S][3]        (empty when not recorded)
    int64_t newton_iters = 0, newton_fail = 0;
};

struct GroupBase {
    int64_t n_done = 0, n_total = 0, launches = 0;
    virtual ~GroupBase() = default;
    virtual void enqueue(int64_t upto, hipStream_t s) = 0;
    virtual void state(std::vector<double> &out) = 0;  // [slots][3]
    virtual const std::vector<int64_t> &slots() const = 0;
    virtual GroupOut collect(hipStream_t s) = 0;
};

// Cluster packing: members of a group sorted by size (largest first), each
// placed in the first block with room; a member never straddles blocks.
// Returns slot -> particle (-1 padding), and per slot the block-relative
// start lane of its member and the member size.
void pack_clusters(const std::vector<std::pair<int64_t, int64_t>> &members, int block,
                   std::vector<int64_t> &slot_particle, std::vector<int> &cbase,
                   std::vector<int> &csize)
{
    std::vector<size_t> order(members.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        return members[x].second - members[x].first > members[y].second - members[y].first;
    });
    std::vector<int> fill;                       // lanes used per block
    std::vector<std::vector<size_t>> content;    // members per block
    size_t first_open = 0;
    for (size_t oi : order) {
        const int sz = (int)(members[oi].second - members[oi].first);
        size_t bi = first_open;
        while (bi < fill.size() && fill[bi] + sz > block) ++bi;
        if (bi == fill.size()) {
            fill.push_back(0);
            content.emplace_back();
        }
        fill[bi] += sz;
        content[bi].push_back(oi);
        while (first_open < fill.size() && fill[first_open] == block) ++first_open;
    }
    const size_t nslots = fill.size() * (size_t)block;
    slot_particle.assign(nslots, -1);
    cbase.assign(nslots, 0);
    csize.assign(nslots, 0);
    for (size_t bi = 0; bi < content.size(); ++bi) {
        int lane = 0;
        for (size_t oi : content[bi]) {
            const int sz = (int)(members[oi].second - members[oi].first);
            for (int j = 0; j < sz; ++j) {
                const size_t slot = bi * (size_t)block + (size_t)(lane + j);
                slot_particle[slot] = members[oi].first + j;
                cbase[slot] = lane;
                csize[slot] = sz;
            }
            lane += sz;
        }
    }
}

template <typename T>
struct GroupRun : GroupBase {
    int n = 0, grid = 0, block = mp::kBlock, n_waves = 0, scheme = 0, shape_kind = -1;
    int max_iter = 1000;
    double eps = 1e-9;
    bool array_noise = false, want_traj = false, want_ens = false, cluster = false;
    bool interactions = true;  // clusters: dipolar sums (the implicit scheme may run without)
    int two_wave = 0;          // Heun kernel: 0 one wave, 1/2 two waves (two_wave_mode)
    int lanes = 1;             // lanes per cluster slot (1 or 2)
    int64_t S = 0, s_next = 0, spl = 1;
    std::vector<int64_t> sched, slot_particle;
    DevBuf<T> d_ex, d_ey, d_ez, d_k, d_al, d_dt, d_sig, d_sr, d_happ, d_mx, d_my, d_mz, d_samples;
    DevBuf<T> d_px, d_py, d_pz, d_w;
    DevBuf<double> d_fred, d_fdt, d_dtd, d_msw, d_noise, d_partial, d_ens;
    DevBuf<int> d_shape, d_cbase, d_csize;
    DevBuf<uint32_t> d_key0, d_key1, d_pid, d_status;
    DevBuf<int64_t> d_sched;
    DevBuf<unsigned long long> d_it, d_fl;
    DevBuf<float> d_htab;      // uniform applied field of one launch (FP32 Heun)
    bool uniform_field = false;
    int field_shape0 = 0;
    double happ0 = 0.0, fdt0 = 0.0;
    mp::IndepArgs<T> a{};

    // cluster_block == 0: one slot per particle of the group (independent
    // particles); otherwise members packed into blocks of cluster_block lanes,
    // cluster_lanes (1 or 2) lanes per slot.
    GroupRun(const simulation::EnsembleSpec &sp, const Plan &plan, int gi, bool traj, bool ens,
             int cluster_block, int cluster_lanes = 1)
    {
        const int64_t b = plan.group_begin[(size_t)gi], e = plan.group_begin[(size_t)gi + 1];
        const int64_t P = sp.member_offsets[sp.n_members];
        sched = plan.sched[(size_t)gi];
        S = (int64_t)sched.size();
        n_total = plan.group_steps[(size_t)gi];
        scheme = sp.scheme;
        eps = sp.implicit_tol;
        max_iter = sp.implicit_max_iter > 0 ? sp.implicit_max_iter : 1000;
        array_noise = sp.noise == MP_NOISE_ARRAY;
        interactions = sp.interactions != 0;
        want_traj = traj;
        want_ens = ens;
        spl = steps_per_launch(sp.scheme, sp.precision);
        n_done = std::min(sp.start_step, n_total);
        cluster = cluster_block > 0;
        std::vector<int> cbase, csize;
        if (cluster) {
            std::vector<std::pair<int64_t, int64_t>> members;  // particle ranges
            for (int64_t i = b; i < e;) {
                const int64_t m = plan.member_of[(size_t)plan.perm[(size_t)i]];
                const int64_t len = sp.member_offsets[m + 1] - sp.member_offsets[m];
                members.emplace_back(sp.member_offsets[m], sp.member_offsets[m] + len);
                i += len;
            }
            block = cluster_block;
            lanes = cluster_lanes;
            pack_clusters(members, block / lanes, slot_particle, cbase, csize);
        } else {
            slot_particle.assign(plan.perm.begin() + b, plan.perm.begin() + e);
        }
        n = (int)slot_particle.size();

        std::vector<double> ex(n, 0.0), ey(n, 0.0), ez(n, 1.0), k(n, 0.0), al(n, 0.0), dt(n, 0.0),
            sig(n, 0.0), sr(n, 0.0), happ(n, 0.0), fred(n, 0.0), fdt(n, 0.0), msw(n, 0.0),
            mx(n, 0.0), my(n, 0.0), mz(n, 1.0), px(n, 0.0), py(n, 0.0), pz(n, 0.0), w(n, 0.0);
        std::vector<int> shape(n, MP_FIELD_CONSTANT);
        std::vector<uint32_t> key0(n, 0), key1(n, 0), pid(n, 0);
        bool first = true;
        for (int i = 0; i < n; ++i) {
            const int64_t p = slot_particle[(size_t)i];
            if (p < 0) continue;
            const int64_t m = plan.member_of[(size_t)p];
            const simulation::ReducedMember &r = plan.red[(size_t)m];
            const int64_t lp = p - sp.member_offsets[m];
            ex[i] = sp.anisotropy_axis[3 * p];
            ey[i] = sp.anisotropy_axis[3 * p + 1];
            ez[i] = sp.anisotropy_axis[3 * p + 2];
            const double *m0 = sp.start_state ? sp.start_state : sp.magnetisation_direction;
            mx[i] = m0[3 * p];
            my[i] = m0[3 * p + 1];
            mz[i] = m0[3 * p + 2];
            k[i] = r.k_red[(size_t)lp];
            al[i] = sp.damping[m];
            dt[i] = r.dt;
            sr[i] = r.sr[(size_t)lp];
            sig[i] = r.sr[(size_t)lp] * std::sqrt(r.dt);
            happ[i] = r.happ;
            fred[i] = r.fapp;
            fdt[i] = r.fapp * r.dt;
            shape[i] = sp.field_shape[m];
            msw[i] = sp.magnetisation[m];
            key0[i] = (uint32_t)(uint64_t)sp.seed[m];
            key1[i] = (uint32_t)((uint64_t)sp.seed[m] >> 32);
            pid[i] = (uint32_t)lp;
            if (cluster) {
                // lib/simulation.cpp:577-583 and lib/field.cpp:212-215
                const double vroot = std::pow(r.average_volume, 1. / 3);
                px[i] = sp.location[3 * p] / vroot;
                py[i] = sp.location[3 * p + 1] / vroot;
                pz[i] = sp.location[3 * p + 2] / vroot;
                const double pre = kMU0 * sp.magnetisation[m] * sp.magnetisation[m] / 8.0 /
                                   kPiHost / r.average_anisotropy;
                w[i] = pre * r.v_red[(size_t)lp];
            }
            if (first) {
                shape_kind = shape[i];
                field_shape0 = shape[i];
                happ0 = happ[i];
                fdt0 = fdt[i];
                first = false;
            } else if (shape_kind != shape[i]) {
                shape_kind = (int)mp::kMixed;
            }
        }
        // one shared field table when every particle sees the same applied field
        uniform_field = sizeof(T) == 4 && scheme == MP_SCHEME_HEUN && !first;
        for (int i = 0; i < n && uniform_field; ++i)
            if (slot_particle[(size_t)i] >= 0)
                uniform_field = shape[i] == field_shape0 && happ[i] == happ0 && fdt[i] == fdt0;
        if (uniform_field) {
            // groups of four aligned below n0, one entry past the last step, and
            // the next group's prefetch: 4 * groups + 5 <= spl + 12
            d_htab.alloc((size_t)spl + 16);
        }
        d_ex.upload(cast_vec<T>(ex));
        d_ey.upload(cast_vec<T>(ey));
        d_ez.upload(cast_vec<T>(ez));
        d_k.upload(cast_vec<T>(k));
        d_al.upload(cast_vec<T>(al));
        d_dt.upload(cast_vec<T>(dt));
        d_sig.upload(cast_vec<T>(sig));
        d_sr.upload(cast_vec<T>(sr));
        d_happ.upload(cast_vec<T>(happ));
        d_mx.upload(cast_vec<T>(mx));
        d_my.upload(cast_vec<T>(my));
        d_mz.upload(cast_vec<T>(mz));
        d_fred.upload(fred);
        d_fdt.upload(fdt);
        d_dtd.upload(dt);
        d_msw.upload(msw);
        d_shape.upload(shape);
        d_key0.upload(key0);
        d_key1.upload(key1);
        d_pid.upload(pid);
        d_status.upload(std::vector<uint32_t>((size_t)n, 0));
        d_sched.upload(sched);
        d_it.upload(std::vector<unsigned long long>((size_t)n, 0));
        d_fl.upload(std::vector<unsigned long long>((size_t)n, 0));
        if (cluster) {
            d_px.upload(cast_vec<T>(px));
            d_py.upload(cast_vec<T>(py));
            d_pz.upload(cast_vec<T>(pz));
            d_w.upload(cast_vec<T>(w));
            d_cbase.upload(cbase);
            d_csize.upload(csize);
        }
        // cluster slots come in whole blocks of block / lanes
        grid = cluster ? n / (block / lanes) : (n + block - 1) / block;
        n_waves = (grid * block) / mp::kWave;
        two_wave = !cluster && !array_noise && scheme == MP_SCHEME_HEUN
                       ? two_wave_mode(n_waves, sizeof(T) == 4)
                       : 0;
        if (want_traj) d_samples.alloc((size_t)S * 3 * (size_t)n);
        if (want_ens) {
            // each recorded sample writes all of its (component, wave) slots;
            // slots of samples taken at step 0 are replaced on the host
            d_partial.alloc((size_t)S * 3 * (size_t)n_waves);
            d_ens.alloc((size_t)S * 3);
        }
        if (array_noise) {
            std::vector<double> nz((size_t)n_total * (size_t)n * 3, 0.0);
            for (int64_t st = 0; st < n_total; ++st)
                for (int i = 0; i < n; ++i) {
                    const int64_t p = slot_particle[(size_t)i];
                    if (p < 0) continue;
                    for (int d = 0; d < 3; ++d)
                        nz[((size_t)st * n + i) * 3 + d] = sp.noise_array[((size_t)st * P + p) * 3 + d];
                }
            d_noise.upload(nz);
        }
        a.n = n;
        a.n_waves = n_waves;
        a.ex = d_ex.p; a.ey = d_ey.p; a.ez = d_ez.p;
        a.k = d_k.p; a.alpha = d_al.p; a.dt = d_dt.p; a.sig = d_sig.p; a.sr = d_sr.p;
        a.happ = d_happ.p; a.fred = d_fred.p; a.fdt = d_fdt.p; a.dtd = d_dtd.p;
        a.shape = d_shape.p;
        a.key0 = d_key0.p; a.key1 = d_key1.p; a.pid = d_pid.p;
        a.msw = d_msw.p;
        a.mx = d_mx.p; a.my = d_my.p; a.mz = d_mz.p;
        a.status = d_status.p;
        a.sched = d_sched.p;
        a.S = S;
        a.samples = want_traj ? d_samples.p : nullptr;
        a.ens_partial = want_ens ? d_partial.p : nullptr;
        a.tag = sp.noise_tag;
        a.renorm = sp.renorm ? 1 : 0;
        a.noise = array_noise ? d_noise.p : nullptr;
    }

    const std::vector<int64_t> &slots() const override { return slot_particle; }

    void launch_cluster(hipStream_t s, const float *htab)
    {
        if constexpr (sizeof(T) == 8) {
            if (scheme == MP_SCHEME_IMPLICIT_MIDPOINT) {
                if (array_noise) launch_implicit_cluster<true>(s);
                else launch_implicit_cluster<false>(s);
                return;
            }
        }
        if (array_noise) launch_cluster_lanes<1, true>(s, htab);  // one lane per slot
        else if (lanes == 2) launch_cluster_lanes<2, false>(s, htab);
        else launch_cluster_lanes<1, false>(s, htab);
    }

    // Partners whose geometry the two-lane FP32 kernel caches per lane (in
    // workgroups of at most 256 lanes): 16, or 0 with MAGPY_AMD_CLUSTER_CACHE=0.
    // Results do not depend on it.
    int cluster_cache() const
    {
        if (sizeof(T) != 4 || lanes != 2 || block > 256 || array_noise) return 0;
        if (const char *env = std::getenv("MAGPY_AMD_CLUSTER_CACHE"))
            if (std::atoi(env) == 0) return 0;
        return 16;
    }

    template <bool kArray>
    void launch_implicit_cluster(hipStream_t s)
    {
        mp::ImplicitClusterArgs ia{
            mp::ClusterArgs<double>{a, d_px.p, d_py.p, d_pz.p, d_w.p, d_cbase.p, d_csize.p}, eps,
            max_iter, d_it.p, d_fl.p, interactions ? 1 : 0};
        const size_t shmem = 3 * (size_t)block * sizeof(mp::Vec4<double>);
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, s, ia);
        };
        switch (shape_kind) {
        case mp::kSine: go(mp::implicit_cluster_kernel<mp::kSine, kArray>); break;
        case mp::kSquare: go(mp::implicit_cluster_kernel<mp::kSquare, kArray>); break;
        case mp::kConstant: go(mp::implicit_cluster_kernel<mp::kConstant, kArray>); break;
        default: go(mp::implicit_cluster_kernel<mp::kMixed, kArray>); break;
        }
    }

    template <int L, bool kArray>
    void launch_cluster_lanes(hipStream_t s, const float *htab)
    {
        if constexpr (L == 2 && !kArray && sizeof(T) == 4) {
            if (cluster_cache() == 16) {
                if (cluster_lds_cache()) launch_cluster_variant<L, kArray, 16, 16>(s, htab);
                else launch_cluster_variant<L, kArray, 16, 0>(s, htab);
                return;
            }
        }
        launch_cluster_variant<L, kArray, 0, 0>(s, htab);
    }

    // the 16 partners after the register-cached ones keep their weights in
    // LDS (MAGPY_AMD_CLUSTER_LDS_CACHE=0: recomputed, the same bits)
    static bool cluster_lds_cache()
    {
        const char *env = std::getenv("MAGPY_AMD_CLUSTER_LDS_CACHE");
        return !(env != nullptr && std::atoi(env) == 0);
    }

    template <int L, bool kArray, int KC, int KL>
    void launch_cluster_variant(hipStream_t s, const float *htab)
    {
        mp::ClusterArgs<T> ca{a, d_px.p, d_py.p, d_pz.p, d_w.p, d_cbase.p, d_csize.p};
        const size_t shmem = 3 * (size_t)(block / L) * sizeof(mp::Vec4<T>) +
                             (size_t)KL * (size_t)block * sizeof(float2);
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, s, ca, htab);
        };
        const bool rn = a.renorm != 0;
        if (htab) {
            if (rn) go(mp::heun_cluster_kernel<T, mp::kConstant, true, true, L, kArray, KC, KL>);
            else go(mp::heun_cluster_kernel<T, mp::kConstant, true, false, L, kArray, KC, KL>);
            return;
        }
        switch (shape_kind) {
        case mp::kSine:
            if (rn) go(mp::heun_cluster_kernel<T, mp::kSine, false, true, L, kArray, KC, KL>);
            else go(mp::heun_cluster_kernel<T, mp::kSine, false, false, L, kArray, KC, KL>);
            break;
        case mp::kSquare:
            if (rn) go(mp::heun_cluster_kernel<T, mp::kSquare, false, true, L, kArray, KC, KL>);
            else go(mp::heun_cluster_kernel<T, mp::kSquare, false, false, L, kArray, KC, KL>);
            break;
        case mp::kConstant:
            if (rn) go(mp::heun_cluster_kernel<T, mp::kConstant, false, true, L, kArray, KC, KL>);
            else go(mp::heun_cluster_kernel<T, mp::kConstant, false, false, L, kArray, KC, KL>);
            break;
        default:
            if (rn) go(mp::heun_cluster_kernel<T, mp::kMixed, false, true, L, kArray, KC, KL>);
            else go(mp::heun_cluster_kernel<T, mp::kMixed, false, false, L, kArray, KC, KL>);
            break;
        }
    }

    void enqueue(int64_t upto, hipStream_t s) override
    {
        upto = std::min(upto, n_total);
        while (n_done < upto) {
            const int64_t n1 = std::min(upto, n_done + spl);
            while (s_next < S && sched[(size_t)s_next] < n_done + 1) ++s_next;
            a.n0 = n_done;
            a.n1 = n1;
            a.s_next = s_next;
            const float *htab = nullptr;
            if (uniform_field) {
                const int64_t g0 = n_done >> 2, g1 = (n1 + 3) >> 2;
                const int64_t count = 4 * (g1 - g0) + 5;
                require((size_t)count <= d_htab.n, "field table too small");
                hipLaunchKernelGGL(mp::field_table_kernel, dim3((unsigned)((count + 255) / 256)),
                                   dim3(256), 0, s, field_shape0, happ0, fdt0, 4 * g0, count,
                                   d_htab.p);
                MP_HIP(hipGetLastError());
                htab = d_htab.p;
            }
            if (cluster)
                launch_cluster(s, htab);
            else if (array_noise)
                launch_shape<T, true>(shape_kind, a, scheme, eps, max_iter, d_it.p, d_fl.p, grid, s,
                                      htab, false);
            else
                launch_shape<T, false>(shape_kind, a, scheme, eps, max_iter, d_it.p, d_fl.p, grid, s,
                                       htab, two_wave);
            MP_HIP(hipGetLastError());
            ++launches;
            n_done = n1;
        }
    }

    void state(std::vector<double> &out) override
    {
        std::vector<T> hx((size_t)n), hy((size_t)n), hz((size_t)n);
        MP_HIP(hipMemcpy(hx.data(), d_mx.p, n * sizeof(T), hipMemcpyDeviceToHost));
        MP_HIP(hipMemcpy(hy.data(), d_my.p, n * sizeof(T), hipMemcpyDeviceToHost));
        MP_HIP(hipMemcpy(hz.data(), d_mz.p, n * sizeof(T), hipMemcpyDeviceToHost));
        out.resize((size_t)n * 3);
        for (int i = 0; i < n; ++i) {
            out[(size_t)i * 3] = (double)hx[(size_t)i];
            out[(size_t)i * 3 + 1] = (double)hy[(size_t)i];
            out[(size_t)i * 3 + 2] = (double)hz[(size_t)i];
        }
    }

    GroupOut collect(hipStream_t s) override
    {
        GroupOut o;
        o.particle_of = slot_particle;
        if (want_ens) {
            hipLaunchKernelGGL(mp::ens_reduce_kernel, dim3((unsigned)(S * 3)), dim3(mp::kBlock), 0,
                               s, d_partial.p, n_waves, S * 3, d_ens.p);
            MP_HIP(hipGetLastError());
        }
        MP_HIP(hipStreamSynchronize(s));
        state(o.state);
        o.status.resize((size_t)n);
        MP_HIP(hipMemcpy(o.status.data(), d_status.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        std::vector<unsigned long long> it((size_t)n), fl((size_t)n);
        MP_HIP(hipMemcpy(it.data(), d_it.p, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        MP_HIP(hipMemcpy(fl.data(), d_fl.p, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            o.newton_iters += (int64_t)it[(size_t)i];
            o.newton_fail += (int64_t)fl[(size_t)i];
        }
        if (want_traj) {
            std::vector<T> hs(d_samples.n);
            if (!hs.empty())
                MP_HIP(hipMemcpy(hs.data(), d_samples.p, hs.size() * sizeof(T), hipMemcpyDeviceToHost));
            o.samples.assign(hs.begin(), hs.end());
        }
        if (want_ens) {
            o.ens.resize((size_t)S * 3);
            MP_HIP(hipMemcpy(o.ens.data(), d_ens.p, o.ens.size() * sizeof(double),
                             hipMemcpyDeviceToHost));
        }
        return o;
    }
};

// ------------------------------------------------------------------ replicas
int pad_to(int64_t x, int64_t q) { return (int)((x + q - 1) / q * q); }

// Dipole tensor of one cluster in reduced units (lib/field.cpp:187-225,
// lib/simulation.cpp:577-583): D[(i,a)][(j,b)] = w_j r^-3 (3 u_a u_b - d_ab),
// w_j = mu0 Ms^2 / (8 pi K_avg) v_j / V_avg, rows and columns in the
// replica_col order, leading dimension ld, zero padded.
void dipole_tensor(const simulation::ReducedMember &r, const double *loc, double ms, size_t n,
                   int ld, std::vector<double> &D)
{
    const double vroot = std::pow(r.average_volume, 1. / 3);
    const double pre = kMU0 * ms * ms / 8.0 / kPiHost / r.average_anisotropy;
    D.assign((size_t)ld * ld, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            if (i == j) continue;
            double d[3], rr = 0.0;
            for (int a = 0; a < 3; ++a) {
                d[a] = (loc[3 * j + a] - loc[3 * i + a]) / vroot;
                rr += d[a] * d[a];
            }
            rr = std::sqrt(rr);
            const double c = pre * r.v_red[j] / (rr * rr * rr);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    D[(size_t)mp::replica_col((int)i, a) * ld + mp::replica_col((int)j, b)] =
                        c * (3 * (d[a] / rr) * (d[b] / rr) - (a == b ? 1.0 : 0.0));
        }
}

void split_host(const std::vector<double> &x, std::vector<__bf16> &hi, std::vector<__bf16> &lo)
{
    hi.resize(x.size());
    lo.resize(x.size());
    for (size_t i = 0; i < x.size(); ++i) {
        const float f = (float)x[i];
        hi[i] = (__bf16)f;
        lo[i] = (__bf16)(f - (float)hi[i]);
    }
}

// Replica layout sizes: particles padded to a multiple of kParticlePad
// (ld = 3 Np), replicas to a multiple of the 128-row block tile.
int replica_ld(int64_t n) { return 3 * pad_to(n, mp::kParticlePad); }
int replica_rows(int64_t reps) { return pad_to(reps, mp::kBM); }

// Which panels the blocks that one XCD runs share: 1 (default) the row
// panels of the state, so each XCD reads an eighth of it and all of the
// tensor (100 MB of L2 fills per 4096 x 1536 product instead of 209 MB);
// 0 (MAGPY_AMD_REPLICA_PANELS=cols) the column panels of the tensor.
int replica_row_panels()
{
    static const int v = [] {
        const char *e = std::getenv("MAGPY_AMD_REPLICA_PANELS");
        return e != nullptr && std::string(e) == "cols" ? 0 : 1;
    }();
    return v;
}

// The BF16x3 product runs the Heun stages in its epilogue unless
// MAGPY_AMD_REPLICA_FUSED=0 when the run is created (then it writes the
// field and separate stage kernels follow, as for the FP32 product; the two
// give the same bits).
bool replica_fused()
{
    const char *e = std::getenv("MAGPY_AMD_REPLICA_FUSED");
    return !(e != nullptr && std::string(e) == "0");
}

void launch_product(int mode, const float *a32, const float *b32, const __bf16 *b_hi,
                    const __bf16 *b_lo, float *c, int rows, int ld, hipStream_t s)
{
    require(rows % mp::kBM == 0 && ld % mp::kBN == 0 && ld % mp::kGemmTile == 0 &&
                ld % (6 * mp::kBK) == 0,
            "replica product: rows or ld do not fit the tiles");
    if (mode == MP_DIPOLAR_GEMM_F32) {
        const dim3 grid((unsigned)(rows / mp::kGemmTile), (unsigned)(ld / mp::kGemmTile));
        hipLaunchKernelGGL(mp::gemm_f32_kernel, grid, dim3(256), 0, s, a32, b32, c, ld, ld);
    } else {
        const int nrb = rows / mp::kBM, ncb = ld / mp::kBN;
        hipLaunchKernelGGL(mp::gemm_bf16x3_ksplit_kernel<0>, dim3((unsigned)(nrb * ncb)), dim3(512),
                           0, s, a32, b_hi, b_lo, c, ld, ld, nrb, ncb, mp::ReplicaArgs{},
                           (int64_t)0, 0.f, (int64_t)0, 0, replica_row_panels());
    }
}

// All members are one cluster replicated: state [rows][ld] replica-major,
// four launches per Heun step (product, stage 1, product, stage 2); the
// products read the FP32 state and split it to BF16 themselves.
struct ReplicaRun : GroupBase {
    int n = 0, reps = 0, ld = 0, rows = 0, mode = MP_DIPOLAR_GEMM_BF16X3, n_waves = 0;
    bool fused = false;
    int shape = MP_FIELD_CONSTANT;
    double happ = 0.0, fapp = 0.0, dt = 0.0;
    bool want_traj = false, want_ens = false;
    int64_t S = 0, s_next = 0;
    std::vector<int64_t> sched, slot_particle;
    DevBuf<float> d_m, d_mt, d_h, d_nz, d_kex, d_key, d_kez, d_edx, d_edy, d_edz, d_sig, d_b32,
        d_samples;
    DevBuf<__bf16> d_bhi, d_blo;
    DevBuf<uint32_t> d_key0, d_key1, d_status;
    DevBuf<double> d_msw, d_partial, d_ens;
    DevBuf<int64_t> d_sched;
    mp::ReplicaArgs a{};

    ReplicaRun(const simulation::EnsembleSpec &sp, const Plan &plan, bool traj, bool ens, int m)
    {
        mode = m;
        fused = mode == MP_DIPOLAR_GEMM_BF16X3 && replica_fused();
        reps = (int)sp.n_members;
        n = (int)(sp.member_offsets[1] - sp.member_offsets[0]);
        ld = replica_ld(n);
        rows = replica_rows(reps);
        sched = plan.sched[0];
        S = (int64_t)sched.size();
        n_total = plan.group_steps[0];
        want_traj = traj;
        want_ens = ens;
        const simulation::ReducedMember &r = plan.red[0];
        shape = sp.field_shape[0];
        happ = r.happ;
        fapp = r.fapp;
        dt = r.dt;
        slot_particle.resize((size_t)reps * n);
        for (int rr = 0; rr < reps; ++rr)
            for (int p = 0; p < n; ++p)
                slot_particle[(size_t)rr * n + p] = sp.member_offsets[rr] + p;
        // per-particle parameters (identical across replicas)
        std::vector<double> kex(n), key(n), kez(n), edx(n), edy(n), edz(n), sig(n);
        for (int p = 0; p < n; ++p) {
            const double *e = sp.anisotropy_axis + 3 * p;
            kex[p] = r.k_red[p] * e[0];
            key[p] = r.k_red[p] * e[1];
            kez[p] = r.k_red[p] * e[2];
            edx[p] = e[0] * r.dt;
            edy[p] = e[1] * r.dt;
            edz[p] = e[2] * r.dt;
            sig[p] = r.sr[p] * std::sqrt(r.dt);
        }
        d_kex.upload(cast_vec<float>(kex));
        d_key.upload(cast_vec<float>(key));
        d_kez.upload(cast_vec<float>(kez));
        d_edx.upload(cast_vec<float>(edx));
        d_edy.upload(cast_vec<float>(edy));
        d_edz.upload(cast_vec<float>(edz));
        d_sig.upload(cast_vec<float>(sig));
        std::vector<uint32_t> k0(reps), k1(reps);
        std::vector<double> msw(reps);
        for (int rr = 0; rr < reps; ++rr) {
            k0[rr] = (uint32_t)(uint64_t)sp.seed[rr];
            k1[rr] = (uint32_t)((uint64_t)sp.seed[rr] >> 32);
            msw[rr] = sp.magnetisation[rr];
        }
        d_key0.upload(k0);
        d_key1.upload(k1);
        d_msw.upload(msw);
        // dipole tensor and its split
        std::vector<double> D;
        dipole_tensor(r, sp.location, sp.magnetisation[0], (size_t)n, ld, D);
        if (mode == MP_DIPOLAR_GEMM_F32) {
            d_b32.upload(cast_vec<float>(D));
        } else {
            std::vector<__bf16> hi, lo;
            split_host(D, hi, lo);
            d_bhi.upload(hi);
            d_blo.upload(lo);
        }
        // state, replica-major and zero padded
        std::vector<double> m0((size_t)rows * ld, 0.0);
        const double *src = sp.start_state ? sp.start_state : sp.magnetisation_direction;
        for (int rr = 0; rr < reps; ++rr)
            for (int p = 0; p < n; ++p)
                for (int c = 0; c < 3; ++c)
                    m0[(size_t)rr * ld + mp::replica_col(p, c)] =
                        src[3 * (sp.member_offsets[rr] + p) + c];
        n_done = std::min(sp.start_step, n_total);
        d_m.upload(cast_vec<float>(m0));
        d_mt.upload(std::vector<float>(m0.size(), 0.f));
        // the fused product keeps the field on chip and passes the step's
        // increments from stage 1 to stage 2; the separate stages pass the field
        if (fused) d_nz.upload(std::vector<float>(m0.size(), 0.f));
        else d_h.upload(std::vector<float>(m0.size(), 0.f));
        d_status.upload(std::vector<uint32_t>((size_t)reps * n, 0));
        d_sched.upload(sched);
        const int threads = reps * n;
        const int grid = (threads + 255) / 256;
        // the fused BF16x3 stages write one ensemble partial per product wave
        n_waves = fused ? (rows / mp::kBM) * (ld / mp::kBN) * 8 : grid * 4;
        if (want_traj) d_samples.alloc((size_t)S * 3 * (size_t)threads);
        if (want_ens) {
            d_partial.alloc((size_t)S * 3 * (size_t)n_waves);
            d_ens.alloc((size_t)S * 3);
        }
        a.n = n;
        a.reps = reps;
        a.ld = ld;
        a.n_waves = n_waves;
        a.kex = d_kex.p; a.key = d_key.p; a.kez = d_kez.p;
        a.edx = d_edx.p; a.edy = d_edy.p; a.edz = d_edz.p;
        a.sig = d_sig.p;
        a.dt = (float)r.dt;
        a.alpha = (float)sp.damping[0];
        a.key0 = d_key0.p; a.key1 = d_key1.p;
        a.msw = d_msw.p;
        a.tag = sp.noise_tag;
        a.renorm = sp.renorm ? 1 : 0;
        a.m = d_m.p; a.mt = d_mt.p; a.h = d_h.p; a.nz = d_nz.p;
        a.status = d_status.p;
        a.sched = d_sched.p;
        a.S = S;
        a.samples = want_traj ? d_samples.p : nullptr;
        a.ens_partial = want_ens ? d_partial.p : nullptr;
    }

    const std::vector<int64_t> &slots() const override { return slot_particle; }

    void enqueue(int64_t upto, hipStream_t s) override
    {
        upto = std::min(upto, n_total);
        const int threads = reps * n;
        const dim3 grid((unsigned)((threads + 255) / 256));
        while (n_done < upto) {
            const int64_t st = n_done;
            const double t1 = (double)(st + 1) * dt;
            const float h1 = (float)applied_host(shape, happ, fapp, t1);
            const float h2 = (float)applied_host(shape, happ, fapp, t1 + dt);
            while (s_next < S && sched[(size_t)s_next] < st + 1) ++s_next;
            int n_rec = 0;
            while (s_next + n_rec < S && sched[(size_t)(s_next + n_rec)] == st + 1) ++n_rec;
            if (fused) {
                // the stages run in the products' epilogues
                require(rows % mp::kBM == 0 && ld % mp::kBN == 0 && ld % (6 * mp::kBK) == 0,
                        "replica product: rows or ld do not fit the tiles");
                const int nrb = rows / mp::kBM, ncb = ld / mp::kBN;
                const dim3 pg((unsigned)(nrb * ncb)), pb(512);
                hipLaunchKernelGGL(mp::gemm_bf16x3_ksplit_kernel<1>, pg, pb, 0, s, d_m.p, d_bhi.p,
                                   d_blo.p, (float *)nullptr, ld, ld, nrb, ncb, a, st, h1,
                                   (int64_t)0, 0, replica_row_panels());
                hipLaunchKernelGGL(mp::gemm_bf16x3_ksplit_kernel<2>, pg, pb, 0, s, d_mt.p, d_bhi.p,
                                   d_blo.p, (float *)nullptr, ld, ld, nrb, ncb, a, st, h2, s_next,
                                   n_rec, replica_row_panels());
            } else {
                launch_product(mode, d_m.p, d_b32.p, d_bhi.p, d_blo.p, d_h.p, rows, ld, s);
                hipLaunchKernelGGL(mp::replica_stage1_kernel, grid, dim3(256), 0, s, a, st, h1);
                launch_product(mode, d_mt.p, d_b32.p, d_bhi.p, d_blo.p, d_h.p, rows, ld, s);
                hipLaunchKernelGGL(mp::replica_stage2_kernel, grid, dim3(256), 0, s, a, st, h2,
                                   s_next, n_rec);
            }
            MP_HIP(hipGetLastError());
            ++launches;
            n_done = st + 1;
        }
    }

    void state(std::vector<double> &out) override
    {
        std::vector<float> hm((size_t)rows * ld);
        MP_HIP(hipMemcpy(hm.data(), d_m.p, hm.size() * sizeof(float), hipMemcpyDeviceToHost));
        out.resize((size_t)reps * n * 3);
        for (int rr = 0; rr < reps; ++rr)
            for (int p = 0; p < n; ++p)
                for (int c = 0; c < 3; ++c)
                    out[((size_t)rr * n + p) * 3 + c] = hm[(size_t)rr * ld + mp::replica_col(p, c)];
    }

    GroupOut collect(hipStream_t s) override
    {
        GroupOut o;
        o.particle_of = slot_particle;
        if (want_ens) {
            hipLaunchKernelGGL(mp::ens_reduce_kernel, dim3((unsigned)(S * 3)), dim3(mp::kBlock), 0,
                               s, d_partial.p, n_waves, S * 3, d_ens.p);
            MP_HIP(hipGetLastError());
        }
        MP_HIP(hipStreamSynchronize(s));
        state(o.state);
        o.status.resize((size_t)reps * n);
        MP_HIP(hipMemcpy(o.status.data(), d_status.p, o.status.size() * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
        if (want_traj) {
            std::vector<float> hs(d_samples.n);
            if (!hs.empty())
                MP_HIP(hipMemcpy(hs.data(), d_samples.p, hs.size() * sizeof(float),
                                 hipMemcpyDeviceToHost));
            o.samples.assign(hs.begin(), hs.end());
        }
        if (want_ens) {
            o.ens.resize((size_t)S * 3);
            MP_HIP(hipMemcpy(o.ens.data(), d_ens.p, o.ens.size() * sizeof(double),
                             hipMemcpyDeviceToHost));
        }
        return o;
    }
};

// Every member is the same cluster: only seeds and initial states differ.
bool shared_geometry(const simulation::EnsembleSpec &sp)
{
    const int64_t n = sp.member_offsets[1] - sp.member_offsets[0];
    for (int64_t m = 1; m < sp.n_members; ++m) {
        const int64_t b = sp.member_offsets[m];
        if (sp.member_offsets[m + 1] - b != n) return false;
        if (sp.magnetisation[m] != sp.magnetisation[0] || sp.damping[m] != sp.damping[0] ||
            sp.temperature[m] != sp.temperature[0] || sp.field_shape[m] != sp.field_shape[0] ||
            sp.field_amplitude[m] != sp.field_amplitude[0] ||
            sp.field_frequency[m] != sp.field_frequency[0])
            return false;
        for (int64_t p = 0; p < n; ++p) {
            if (sp.radius[b + p] != sp.radius[p] || sp.anisotropy[b + p] != sp.anisotropy[p])
                return false;
            for (int c = 0; c < 3; ++c)
                if (sp.location[3 * (b + p) + c] != sp.location[3 * p + c] ||
                    sp.anisotropy_axis[3 * (b + p) + c] != sp.anisotropy_axis[3 * p + c])
                    return false;
        }
    }
    return true;
}

}  // namespace

struct mp_session {
    Plan plan;
    HostCopy host;
    std::vector<std::unique_ptr<GroupBase>> groups;
    StreamGuard sg;
    // own_stream false: launches go to the null stream and nothing is created
    explicit mp_session(bool own_stream = true) : sg(own_stream) {}
    bool want_traj = false, want_ens = false;
    int device = 0;
    int cluster_block = 0;
    int cluster_lanes = 0;
    int dipolar = 0;
    double kernel_ms = 0.0;
};

namespace simulation {

ReducedMember reduce_member(const double *radius, const double *anisotropy, size_t n,
                            double magnetisation, double damping, double temperature,
                            double time_step, double end_time, double field_amplitude,
                            double field_frequency)
{
#pragma clang fp contract(off)
    ReducedMember r;
    std::vector<double> volume(n);
    for (size_t p = 0; p < n; ++p)
        volume[p] = 4.0 / 3.0 * kPiHost * radius[p] * radius[p] * radius[p];
    double vsum = 0.0;
    for (size_t p = 0; p < n; ++p) vsum += volume[p];
    r.average_volume = vsum / (double)n;
    double ksum = 0.0;
    for (size_t p = 0; p < n; ++p) ksum += anisotropy[p];
    r.average_anisotropy = ksum / (double)n;
    r.v_red.resize(n);
    r.k_red.resize(n);
    r.sr.resize(n);
    for (size_t p = 0; p < n; ++p) {
        r.v_red[p] = volume[p] / r.average_volume;
        r.k_red[p] = anisotropy[p] / r.average_anisotropy;
        r.sr[p] = std::sqrt(damping * kKB * temperature / (r.average_anisotropy * volume[p]) /
                            (1 + damping * damping));
    }
    r.anisotropy_field = 2 * r.average_anisotropy / kMU0 / magnetisation;
    r.time_factor = kGYROMAG * kMU0 * r.anisotropy_field / (1 + damping * damping);
    r.dt = time_step * r.time_factor;
    r.end_time = end_time * r.time_factor;
    r.happ = field_amplitude / r.anisotropy_field;
    r.fapp = field_frequency / r.time_factor;
    return r;
}

int64_t sample_schedule(double dt, double end_time, int64_t max_samples, std::vector<int64_t> &k)
{
    // The reference's loop (lib/simulation.cpp:342-406) steps while
    // t = step*dt <= s*Ts and records the state one step before it stopped;
    // here the first such step is found by search instead of by stepping.
    k.assign((size_t)max_samples, 0);
    if (max_samples <= 1) return 0;
    const double ts = end_time / (double)(max_samples - 1);
    int64_t step = 0;
    double t = 0.0;
    for (int64_t s = 1; s < max_samples; ++s) {
        const double x = (double)s * ts;
        if (t <= x) {
            double guess = std::floor(x / dt);
            int64_t j = guess > 9.0e18 ? INT64_MAX / 2 : (int64_t)guess;
            if (j < step + 1) j = step + 1;
            while ((double)j * dt <= x) ++j;
            while (j - 1 > step && (double)(j - 1) * dt > x) --j;
            step = j;
            t = (double)step * dt;
        }
        k[(size_t)s] = step - 1;
    }
    return step;
}

void validate(const EnsembleSpec &sp)
{
    require(sp.n_members >= 1, "n_members must be >= 1");
    require(sp.member_offsets != nullptr, "member_offsets is required");
    require(sp.member_offsets[0] == 0, "member_offsets[0] must be 0");
    for (int64_t m = 0; m < sp.n_members; ++m)
        require(sp.member_offsets[m + 1] > sp.member_offsets[m],
                "every member needs at least one particle");
    const int64_t P = sp.member_offsets[sp.n_members];
    require(P <= (int64_t)INT32_MAX, "too many particles for one call");
    require(sp.radius && sp.anisotropy && sp.anisotropy_axis && sp.magnetisation_direction &&
                sp.location,
            "particle arrays are required");
    require(sp.magnetisation && sp.damping && sp.temperature && sp.field_shape &&
                sp.field_amplitude && sp.field_frequency && sp.seed,
            "member arrays are required");
    require(sp.max_samples >= 1,
            "max_samples must be >= 1 (-1 never terminates in the reference loop)");
    require(sp.time_step > 0 && std::isfinite(sp.time_step), "time_step must be > 0");
    require(sp.end_time >= 0 && std::isfinite(sp.end_time), "end_time must be >= 0");
    require(sp.scheme == MP_SCHEME_HEUN || sp.scheme == MP_SCHEME_IMPLICIT_MIDPOINT,
            "unknown scheme");
    require(sp.precision == MP_PRECISION_FP32 || sp.precision == MP_PRECISION_FP64,
            "unknown precision");
    require(sp.noise == MP_NOISE_PHILOX || sp.noise == MP_NOISE_ARRAY, "unknown noise source");
    require(sp.dipolar >= MP_DIPOLAR_AUTO && sp.dipolar <= MP_DIPOLAR_GEMM_F32,
            "unknown dipolar mode");
    require(sp.start_step >= 0, "start_step must be >= 0");
    require(sp.noise != MP_NOISE_ARRAY || sp.noise_array != nullptr,
            "noise_array is required for MP_NOISE_ARRAY");
    for (int64_t m = 0; m < sp.n_members; ++m) {
        const int fs = sp.field_shape[m];
        require(fs == MP_FIELD_SINE || fs == MP_FIELD_SQUARE || fs == MP_FIELD_CONSTANT,
                "Must specify valid field::options enum");
        require(sp.magnetisation[m] > 0 && sp.damping[m] >= 0 && sp.temperature[m] >= 0,
                "magnetisation must be > 0, damping and temperature >= 0");
    }
}

}  // namespace simulation

namespace {

// Layout of the interacting-cluster kernel: `block` lanes per workgroup and
// `lanes` (1 or 2) lanes per particle slot, so a workgroup holds
// block / lanes slots, which must take the largest cluster.
struct ClusterLayout {
    int block = 0, lanes = 1;
};

// The smallest power-of-two block of at least one wave whose slots hold the
// largest cluster (0: none up to max_block lanes).
int cluster_min_block(int64_t nmax, int lanes, int max_block = 1024)
{
    for (int b = mp::kWave; b <= max_block; b *= 2)
        if (b / lanes >= nmax) return b;
    return 0;
}

// What the cluster kernels of a run allow: the implicit midpoint kernel and
// caller-given noise run one lane per slot; the implicit kernel's workgroups
// hold at most kImplicitClusterBlock lanes (FP64 Newton state per lane).
struct ClusterLimits {
    bool one_lane = false;
    int max_block = 1024;
};

ClusterLimits cluster_limits(const simulation::EnsembleSpec &sp)
{
    ClusterLimits c;
    c.one_lane = sp.scheme == MP_SCHEME_IMPLICIT_MIDPOINT || sp.noise == MP_NOISE_ARRAY;
    if (sp.scheme == MP_SCHEME_IMPLICIT_MIDPOINT) c.max_block = mp::kImplicitClusterBlock;
    return c;
}

// Two lanes per particle pay while one lane per particle would leave fewer
// than this many waves per CU: the pair sums of a slot run in half the time
// and the extra waves hide the LDS latency of the pair loop.
constexpr int kSplitBelowWavesPerCU = 16;

// Lanes per slot forced by MAGPY_AMD_CLUSTER_LANES (1 or 2), else 0.
int env_cluster_lanes()
{
    const char *env = std::getenv("MAGPY_AMD_CLUSTER_LANES");
    if (env == nullptr) return 0;
    return std::atoi(env) == 2 ? 2 : 1;
}

// Untimed choice of the layout: MAGPY_AMD_CLUSTER_LANES and
// MAGPY_AMD_CLUSTER_BLOCK (lanes per workgroup) when set and valid, else two
// lanes per slot for ensembles of few waves, and the smallest block.
ClusterLayout cluster_layout(int64_t nmax, int64_t particles, ClusterLimits lim = {})
{
    ClusterLayout l;
    if (lim.one_lane) {
        l.lanes = 1;
    } else if (const int forced = env_cluster_lanes()) {
        l.lanes = forced;
    } else {
        int dev = 0, cus = 0;
        MP_HIP(hipGetDevice(&dev));
        MP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        l.lanes = particles < (int64_t)kSplitBelowWavesPerCU * mp::kWave * cus ? 2 : 1;
    }
    if (cluster_min_block(nmax, l.lanes, lim.max_block) == 0) l.lanes = 1;
    if (const char *env = std::getenv("MAGPY_AMD_CLUSTER_BLOCK")) {
        const int v = std::atoi(env);
        if (v / l.lanes >= nmax && v % mp::kWave == 0 && v <= lim.max_block) {
            l.block = v;
            return l;
        }
    }
    l.block = cluster_min_block(nmax, l.lanes, lim.max_block);
    return l;
}

// Autotuner of the cluster layout: for runs long enough to repay it, time a
// few trial steps of the first schedule group with every candidate (one and
// two lanes per slot, unless MAGPY_AMD_CLUSTER_LANES fixes it, each with the
// power-of-two blocks from the smallest that holds the largest cluster up to
// 1024 lanes) on throw-away copies of the state, and keep the fastest.
// Results are cached per (largest cluster, particles, precision) for the
// life of the process.  Trials run from the group's own first step (a
// resumed run starts at start_step), and a group too short to time them
// keeps the untimed layout, uncached.
ClusterLayout autotune_cluster_layout(const simulation::EnsembleSpec &sp, const Plan &plan,
                                      int64_t nmax, hipStream_t s)
{
    static std::map<std::tuple<int64_t, int64_t, int, int>, ClusterLayout> cache;
    const int forced = env_cluster_lanes();
    const auto key = std::make_tuple(nmax, (int64_t)plan.perm.size(), (int)sp.precision, forced);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    const int64_t trial_steps = 64;
    const int64_t first = std::min(sp.start_step, plan.group_steps[0]);
    if (plan.group_steps[0] - first < 4 + trial_steps)
        return cluster_layout(nmax, (int64_t)plan.perm.size());
    hipEvent_t e0, e1;
    MP_HIP(hipEventCreate(&e0));
    MP_HIP(hipEventCreate(&e1));
    ClusterLayout best;
    float best_ms = 0.f;
    for (int lanes = 1; lanes <= 2; ++lanes) {
        if (forced != 0 && lanes != forced) continue;
        const int b0 = cluster_min_block(nmax, lanes);
        for (int b = b0; b != 0 && b <= 1024; b *= 2) {
            std::unique_ptr<GroupBase> g;
            if (sp.precision == MP_PRECISION_FP32)
                g.reset(new GroupRun<float>(sp, plan, 0, false, false, b, lanes));
            else
                g.reset(new GroupRun<double>(sp, plan, 0, false, false, b, lanes));
            g->enqueue(g->n_done + 4, s);  // warm the code object
            MP_HIP(hipEventRecord(e0, s));
            g->enqueue(g->n_done + trial_steps, s);
            MP_HIP(hipEventRecord(e1, s));
            MP_HIP(hipStreamSynchronize(s));
            float ms = 0.f;
            MP_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (best.block == 0 || ms < best_ms) {
                best.block = b;
                best.lanes = lanes;
                best_ms = ms;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cache[key] = best;
    return best;
}

// With an arena open on this thread (the no-allocation entry point), every
// device buffer of the session comes from it, launches go to the null
// stream, no stream or event is created and the layout is not autotuned.
mp_session *session_create(const simulation::EnsembleSpec &sp, bool traj, bool ens)
{
    const bool in_arena = current_arena() != nullptr;
    simulation::validate(sp);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        throw simulation::Error(MP_ERR_NODEV, "no HIP device visible");
    require(sp.device >= 0 && sp.device < ndev, "device ordinal out of range");
    MP_HIP(hipSetDevice(sp.device));

    std::unique_ptr<mp_session> ss(new mp_session(!in_arena));
    ss->device = sp.device;
    ss->want_traj = traj;
    ss->want_ens = ens;
    Plan &plan = ss->plan;
    const int64_t P = sp.member_offsets[sp.n_members];
    HostCopy &h = ss->host;
    h.M = sp.n_members;
    h.P = P;
    h.S = sp.max_samples;
    const double *m_start = sp.start_state ? sp.start_state : sp.magnetisation_direction;
    h.m0.assign(m_start, m_start + 3 * P);
    h.start = sp.start_step;
    h.ms.assign(sp.magnetisation, sp.magnetisation + sp.n_members);
    h.shape.assign(sp.field_shape, sp.field_shape + sp.n_members);

    // reduce every member, schedule per member, group members by schedule.
    // Every kernel reads its lane's reduced time step, so members whose
    // integer schedules (sample steps and step count) are equal share one
    // launch even when their reduced steps differ (other damping or
    // anisotropy).  FP32 Heun keeps one group per reduced step: there a
    // group whose lanes all see one applied field reads it from a shared
    // table, whose bits differ from the per-lane evaluation, and a member's
    // result must not depend on the rest of the ensemble.
    const bool merge_by_schedule =
        !(sp.precision == MP_PRECISION_FP32 && sp.scheme == MP_SCHEME_HEUN);
    plan.red.reserve((size_t)sp.n_members);
    plan.member_of.resize((size_t)P);
    std::map<std::pair<double, double>, int> step_key;        // (dt, end) -> schedule
    std::vector<std::vector<int64_t>> scheds;                 // schedule + step count
    std::map<std::vector<int64_t>, int> sched_to_group;
    std::vector<int> member_group((size_t)sp.n_members), group_sched;
    for (int64_t m = 0; m < sp.n_members; ++m) {
        const int64_t b = sp.member_offsets[m], e = sp.member_offsets[m + 1];
        plan.red.push_back(simulation::reduce_member(
            sp.radius + b, sp.anisotropy + b, (size_t)(e - b), sp.magnetisation[m], sp.damping[m],
            sp.temperature[m], sp.time_step, sp.end_time, sp.field_amplitude[m],
            sp.field_frequency[m]));
        for (int64_t p = b; p < e; ++p) plan.member_of[(size_t)p] = m;
        const auto key = std::make_pair(plan.red.back().dt, plan.red.back().end_time);
        auto it = step_key.find(key);
        if (it == step_key.end()) {
            std::vector<int64_t> k;
            const int64_t steps =
                simulation::sample_schedule(key.first, key.second, sp.max_samples, k);
            k.push_back(steps);
            it = step_key.emplace(key, (int)scheds.size()).first;
            scheds.push_back(std::move(k));
        }
        const int si = it->second;
        int g = -1;
        if (merge_by_schedule) {
            auto jt = sched_to_group.find(scheds[(size_t)si]);
            if (jt == sched_to_group.end())
                jt = sched_to_group.emplace(scheds[(size_t)si], (int)group_sched.size()).first;
            g = jt->second;
        } else {
            g = si;   // groups are the distinct (dt, end) pairs in order
        }
        if (g == (int)group_sched.size()) group_sched.push_back(si);
        member_group[(size_t)m] = g;
    }
    const int G = (int)group_sched.size();
    plan.group_begin.assign((size_t)G + 1, 0);
    for (int g = 0; g < G; ++g) {
        for (int64_t m = 0; m < sp.n_members; ++m)
            if (member_group[(size_t)m] == g)
                for (int64_t p = sp.member_offsets[m]; p < sp.member_offsets[m + 1]; ++p)
                    plan.perm.push_back(p);
        plan.group_begin[(size_t)g + 1] = (int64_t)plan.perm.size();
        std::vector<int64_t> k = scheds[(size_t)group_sched[(size_t)g]];
        const int64_t steps = k.back();
        k.pop_back();
        plan.sched.push_back(std::move(k));
        plan.group_steps.push_back(steps);
    }
    if (sp.noise == MP_NOISE_ARRAY)
        for (int g = 0; g < G; ++g)
            require(plan.group_steps[(size_t)g] <= sp.noise_steps,
                    "Exceeded allocated random number array size (noise_steps < steps)");
    // interacting clusters: one block holds whole clusters, so the block must
    // be at least the largest cluster.  The implicit scheme couples a
    // cluster's particles through its Newton norms even without interactions
    // (one 3N system per cluster, lib/integrators.cpp:576-651), so it takes
    // the cluster path for any member of several particles.
    ClusterLayout layout;
    int64_t nmax = 1;
    if (sp.interactions || sp.scheme == MP_SCHEME_IMPLICIT_MIDPOINT) {
        for (int64_t m = 0; m < sp.n_members; ++m)
            nmax = std::max(nmax, sp.member_offsets[m + 1] - sp.member_offsets[m]);
        if (nmax > 1) {
            const ClusterLimits lim = cluster_limits(sp);
            layout = cluster_layout(nmax, (int64_t)plan.perm.size(), lim);
            if (layout.block == 0)
                throw simulation::Error(
                    MP_ERR_UNSUPPORTED,
                    sp.scheme == MP_SCHEME_IMPLICIT_MIDPOINT
                        ? "the implicit midpoint scheme supports clusters of at most " +
                              std::to_string(lim.max_block) + " particles"
                        : std::string("clusters of more than 1024 particles are not supported"));
        }
    }
    int cluster_block = layout.block;
    // replicas of one cluster: the dipolar field as one product per stage
    int gemm = 0;
    if (cluster_block != 0 && G == 1 && sp.scheme == MP_SCHEME_HEUN &&
        sp.noise == MP_NOISE_PHILOX && sp.precision == MP_PRECISION_FP32 &&
        sp.dipolar != MP_DIPOLAR_PAIRS && shared_geometry(sp)) {
        const int64_t n = sp.member_offsets[1] - sp.member_offsets[0];
        if (sp.dipolar == MP_DIPOLAR_GEMM_BF16X3 || sp.dipolar == MP_DIPOLAR_GEMM_F32)
            gemm = sp.dipolar;
        else if (sp.n_members >= 64 && n >= 32)
            gemm = MP_DIPOLAR_GEMM_BF16X3;
    }
    if (gemm != 0) {
        ss->cluster_block = 0;
        ss->dipolar = gemm;
        ss->groups.emplace_back(new ReplicaRun(sp, plan, traj, ens, gemm));
        return ss.release();
    }
    if (sp.dipolar == MP_DIPOLAR_GEMM_BF16X3 || sp.dipolar == MP_DIPOLAR_GEMM_F32)
        throw simulation::Error(MP_ERR_UNSUPPORTED,
                                "the dipolar product needs an FP32 Heun ensemble of replicas of "
                                "one cluster with the Philox stream");
    // pair sums: pick the block size by timing the candidates on long runs
    if (cluster_block != 0 && std::getenv("MAGPY_AMD_CLUSTER_BLOCK") == nullptr &&
        sp.noise == MP_NOISE_PHILOX && sp.scheme == MP_SCHEME_HEUN &&
        plan.group_steps[0] >= 20000 && !in_arena)
        layout = autotune_cluster_layout(sp, plan, nmax, ss->sg.s);
    cluster_block = layout.block;
    ss->dipolar = cluster_block && sp.interactions ? MP_DIPOLAR_PAIRS : 0;
    ss->cluster_block = cluster_block;
    ss->cluster_lanes = cluster_block ? layout.lanes : 0;
    for (int g = 0; g < G; ++g) {
        if (sp.precision == MP_PRECISION_FP32 && sp.scheme == MP_SCHEME_HEUN)
            ss->groups.emplace_back(
                new GroupRun<float>(sp, plan, g, traj, ens, cluster_block, layout.lanes));
        else
            ss->groups.emplace_back(
                new GroupRun<double>(sp, plan, g, traj, ens, cluster_block, layout.lanes));
    }
    return ss.release();
}

void session_advance(mp_session *ss, int64_t n)
{
    MP_HIP(hipSetDevice(ss->device));
    if (ss->sg.e0 == nullptr) {
        // no events of its own: the host's clock around the synchronised block
        const auto t0 = std::chrono::steady_clock::now();
        for (auto &g : ss->groups) g->enqueue(g->n_done + n, ss->sg.s);
        MP_HIP(hipStreamSynchronize(ss->sg.s));
        ss->kernel_ms +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                .count();
        return;
    }
    MP_HIP(hipEventRecord(ss->sg.e0, ss->sg.s));
    for (auto &g : ss->groups) g->enqueue(g->n_done + n, ss->sg.s);
    MP_HIP(hipEventRecord(ss->sg.e1, ss->sg.s));
    MP_HIP(hipStreamSynchronize(ss->sg.s));
    float ms = 0.f;
    MP_HIP(hipEventElapsedTime(&ms, ss->sg.e0, ss->sg.e1));
    ss->kernel_ms += ms;
}

void session_finish(mp_session *ss, mp_ensemble_out &out)
{
    int64_t remaining = 0;
    for (auto &g : ss->groups) remaining = std::max(remaining, g->n_total - g->n_done);
    if (remaining > 0) session_advance(ss, remaining);
    MP_HIP(hipSetDevice(ss->device));
    const HostCopy &h = ss->host;
    const Plan &plan = ss->plan;
    const int64_t P = h.P, S = h.S;
    const bool want_traj = ss->want_traj && (out.mx || out.my || out.mz);
    std::vector<double> state((size_t)P * 3), traj, ens;
    std::vector<uint32_t> status((size_t)P, 0);
    if (want_traj) traj.assign((size_t)S * 3 * (size_t)P, 0.0);
    if (out.ens_m && ss->want_ens) ens.assign((size_t)S * 3, 0.0);
    int64_t newton_iters = 0, newton_fail = 0;
    for (size_t gi = 0; gi < ss->groups.size(); ++gi) {
        GroupOut o = ss->groups[gi]->collect(ss->sg.s);
        const int n = (int)o.particle_of.size();
        const std::vector<int64_t> &sched = plan.sched[gi];
        for (int i = 0; i < n; ++i) {
            const int64_t p = o.particle_of[(size_t)i];
            if (p < 0) continue;
            for (int d = 0; d < 3; ++d) state[(size_t)p * 3 + d] = o.state[(size_t)i * 3 + d];
            status[(size_t)p] = o.status[(size_t)i];
        }
        newton_iters += o.newton_iters;
        newton_fail += o.newton_fail;
        // samples recorded at the start step are the start state exactly as
        // given; samples before a resumed run's start step are not known
        const double nan = std::numeric_limits<double>::quiet_NaN();
        if (want_traj)
            for (int64_t s = 0; s < S; ++s)
                for (int i = 0; i < n; ++i) {
                    const int64_t p = o.particle_of[(size_t)i];
                    if (p < 0) continue;
                    for (int d = 0; d < 3; ++d)
                        traj[((size_t)s * 3 + d) * (size_t)P + (size_t)p] =
                            sched[(size_t)s] < h.start ? nan
                            : sched[(size_t)s] == h.start
                                ? h.m0[(size_t)p * 3 + d]
                                : o.samples[((size_t)s * 3 + d) * (size_t)n + (size_t)i];
                }
        if (!ens.empty())
            for (int64_t s = 0; s < S; ++s)
                for (int d = 0; d < 3; ++d) {
                    double v = o.ens[(size_t)s * 3 + d];
                    if (sched[(size_t)s] < h.start) {
                        v = nan;
                    } else if (sched[(size_t)s] == h.start) {
                        v = 0.0;
                        for (int i = 0; i < n; ++i) {
                            const int64_t p = o.particle_of[(size_t)i];
                            if (p < 0) continue;
                            v += h.m0[(size_t)p * 3 + d] * h.ms[(size_t)plan.member_of[(size_t)p]];
                        }
                    }
                    ens[(size_t)s * 3 + d] += v;
                }
    }

    // ---- results in SI units (lib/simulation.cpp:610-621)
    auto fill_time_field = [&](int64_t m, double *time, double *field) {
        const simulation::ReducedMember &r = plan.red[(size_t)m];
        const double ts = S > 1 ? r.end_time / (double)(S - 1) : 0.0;
        for (int64_t s = 0; s < S; ++s) {
            const double tr = s == 0 ? 0.0 : (double)s * ts;
            if (time) time[s] = tr / r.time_factor;
            if (field)
                field[s] = applied_host(h.shape[(size_t)m], r.happ, r.fapp, tr) * r.anisotropy_field;
        }
    };
    fill_time_field(0, out.time, out.field);
    if (out.member_time || out.member_field)
        for (int64_t m = 0; m < h.M; ++m)
            fill_time_field(m, out.member_time ? out.member_time + m * S : nullptr,
                            out.member_field ? out.member_field + m * S : nullptr);
    if (want_traj) {
        double *dst[3] = {out.mx, out.my, out.mz};
        for (int64_t p = 0; p < P; ++p) {
            const double ms = h.ms[(size_t)plan.member_of[(size_t)p]];
            for (int d = 0; d < 3; ++d) {
                if (!dst[d]) continue;
                for (int64_t s = 0; s < S; ++s)
                    dst[d][p * S + s] = traj[((size_t)s * 3 + d) * (size_t)P + (size_t)p] * ms;
            }
        }
    }
    if (out.ens_m && !ens.empty()) std::memcpy(out.ens_m, ens.data(), ens.size() * sizeof(double));
    if (out.final_state) std::memcpy(out.final_state, state.data(), state.size() * sizeof(double));
    if (out.status) std::memcpy(out.status, status.data(), status.size() * sizeof(uint32_t));
    out.n_steps = plan.group_steps[0];
    out.newton_iters = newton_iters;
    out.newton_failures = newton_fail;
    out.kernel_ms = ss->kernel_ms;
}

}  // namespace

namespace simulation {

void ensemble_run(const EnsembleSpec &sp, EnsembleResult &out)
{
    std::unique_ptr<mp_session> ss(
        session_create(sp, out.mx || out.my || out.mz, out.ens_m != nullptr));
    session_finish(ss.get(), out);
}

}  // namespace simulation

// ============================================ fixed scratch (no allocation)
// Device memory of mp_ensemble_run_noalloc: one array per GPU, placed when
// the library's code object loads on that GPU and held until the process
// ends.  Host code never uses the name `scratch` as an address (on the host
// it is only a stand-in for the symbol): the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// The calling thread's current device's copy of `scratch`.
void *scratch_on(int device)
{
    require(device >= 0 && device < kMaxDevices, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device] == nullptr) {
        void *ptr = nullptr;
        MP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        require(ptr != nullptr, "no device address for the scratch array");
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int mp_ensemble_run(const mp_ensemble_spec *spec, mp_ensemble_out *out)
{
    return guarded([&] {
        if (!spec || !out) throw simulation::Error(MP_ERR_ARG, "spec and out are required");
        simulation::ensemble_run(*spec, *out);
    });
}

int mp_ensemble_run_noalloc(const mp_ensemble_spec *spec, mp_ensemble_out *out)
{
    return guarded([&] {
        if (!spec || !out) throw simulation::Error(MP_ERR_ARG, "spec and out are required");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw simulation::Error(MP_ERR_NODEV, "no HIP device visible");
        require(spec->device >= 0 && spec->device < ndev, "device ordinal out of range");
        MP_HIP(hipSetDevice(spec->device));
        Arena arena;
        arena.base = static_cast<char *>(scratch_on(spec->device));
        arena.cap = kScratchBytes;
        arena.who = "mp_ensemble_run_noalloc";
        ArenaScope scope(&arena);
        simulation::ensemble_run(*spec, *out);
    });
}

int mp_full_dynamics(const double *radius, const double *anisotropy,
                     const double *anisotropy_axis, const double *magnetisation_direction,
                     const double *location, size_t n_particles, double magnetisation,
                     double damping, double temperature, int renorm, int interactions,
                     int use_implicit, double eps, double time_step, double end_time,
                     size_t max_samples, long seed, int field_shape, double field_amplitude,
                     double field_frequency, int precision, double *time, double *field,
                     double *mx, double *my, double *mz)
{
    int64_t offs[2] = {0, (int64_t)n_particles};
    int32_t shape = field_shape;
    int64_t sd = seed;
    mp_ensemble_spec sp{};
    sp.n_members = 1;
    sp.member_offsets = offs;
    sp.radius = radius;
    sp.anisotropy = anisotropy;
    sp.anisotropy_axis = anisotropy_axis;
    sp.magnetisation_direction = magnetisation_direction;
    sp.location = location;
    sp.magnetisation = &magnetisation;
    sp.damping = &damping;
    sp.temperature = &temperature;
    sp.field_shape = &shape;
    sp.field_amplitude = &field_amplitude;
    sp.field_frequency = &field_frequency;
    sp.seed = &sd;
    sp.time_step = time_step;
    sp.end_time = end_time;
    sp.max_samples = (int64_t)max_samples;
    sp.renorm = renorm;
    sp.interactions = interactions;
    sp.scheme = use_implicit ? MP_SCHEME_IMPLICIT_MIDPOINT : MP_SCHEME_HEUN;
    sp.precision = precision;
    sp.implicit_tol = eps;
    sp.implicit_max_iter = 1000;
    sp.noise = MP_NOISE_PHILOX;
    mp_ensemble_out o{};
    o.time = time;
    o.field = field;
    o.mx = mx;
    o.my = my;
    o.mz = mz;
    return mp_ensemble_run(&sp, &o);
}

int mp_session_create(const mp_ensemble_spec *spec, int want_trajectories, int want_ensemble,
                      mp_session **session)
{
    return guarded([&] {
        if (!spec || !session) throw simulation::Error(MP_ERR_ARG, "spec and session are required");
        *session = session_create(*spec, want_trajectories != 0, want_ensemble != 0);
    });
}

int mp_session_advance(mp_session *session, int64_t n_steps, int64_t *steps_done)
{
    return guarded([&] {
        if (!session || n_steps < 0) throw simulation::Error(MP_ERR_ARG, "bad session arguments");
        session_advance(session, n_steps);
        if (steps_done) *steps_done = session->groups[0]->n_done;
    });
}

int mp_session_info(mp_session *session, int64_t *n_steps_total, int64_t *launches,
                    double *kernel_ms, int32_t *cluster_block, int32_t *dipolar,
                    int32_t *cluster_lanes)
{
    return guarded([&] {
        if (!session) throw simulation::Error(MP_ERR_ARG, "session is required");
        int64_t tot = 0, l = 0;
        for (auto &g : session->groups) {
            tot = std::max(tot, g->n_total);
            l += g->launches;
        }
        if (n_steps_total) *n_steps_total = tot;
        if (launches) *launches = l;
        if (kernel_ms) *kernel_ms = session->kernel_ms;
        if (cluster_block) *cluster_block = session->cluster_block;
        if (dipolar) *dipolar = session->dipolar;
        if (cluster_lanes) *cluster_lanes = session->cluster_lanes;
    });
}

int mp_session_state(mp_session *session, double *state, int64_t *step)
{
    return guarded([&] {
        if (!session || !state) throw simulation::Error(MP_ERR_ARG, "session and state are required");
        MP_HIP(hipSetDevice(session->device));
        MP_HIP(hipStreamSynchronize(session->sg.s));
        for (size_t gi = 0; gi < session->groups.size(); ++gi) {
            std::vector<double> st;
            session->groups[gi]->state(st);
            const std::vector<int64_t> &slots = session->groups[gi]->slots();
            for (size_t i = 0; i < slots.size(); ++i) {
                const int64_t p = slots[i];
                if (p < 0) continue;
                for (int d = 0; d < 3; ++d) state[p * 3 + d] = st[i * 3 + d];
            }
        }
        if (step) *step = session->groups[0]->n_done;
    });
}

int mp_session_finish(mp_session *session, mp_ensemble_out *out)
{
    return guarded([&] {
        if (!session || !out) throw simulation::Error(MP_ERR_ARG, "session and out are required");
        session_finish(session, *out);
    });
}

void mp_session_destroy(mp_session *session) { delete session; }

int mp_sample_schedule(double dt_reduced, double end_reduced, int64_t max_samples, int64_t *k,
                       int64_t *n_steps)
{
    return guarded([&] {
        if (max_samples < 1 || !k || !n_steps || !(dt_reduced > 0))
            throw simulation::Error(MP_ERR_ARG, "invalid schedule arguments");
        std::vector<int64_t> v;
        *n_steps = simulation::sample_schedule(dt_reduced, end_reduced, max_samples, v);
        std::memcpy(k, v.data(), v.size() * sizeof(int64_t));
    });
}

int mp_dipolar_field(const double *location, const double *radius, const double *anisotropy,
                     size_t n_particles, double magnetisation, const double *m, size_t reps,
                     int mode, double *h, int device)
{
    return guarded([&] {
        if (!location || !radius || !anisotropy || !m || !h || n_particles == 0 || reps == 0 ||
            !(magnetisation > 0))
            throw simulation::Error(MP_ERR_ARG, "invalid dipolar field arguments");
        if (mode != MP_DIPOLAR_PAIRS && mode != MP_DIPOLAR_GEMM_BF16X3 && mode != MP_DIPOLAR_GEMM_F32)
            throw simulation::Error(MP_ERR_ARG, "unknown dipolar mode");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw simulation::Error(MP_ERR_NODEV, "no HIP device visible");
        if (device < 0 || device >= ndev) throw simulation::Error(MP_ERR_ARG, "bad device");
        MP_HIP(hipSetDevice(device));
        const simulation::ReducedMember r = simulation::reduce_member(
            radius, anisotropy, n_particles, magnetisation, 0.1, 0.0, 1.0, 1.0, 0.0, 0.0);
        const int ld = replica_ld((int64_t)n_particles);
        const int rows = replica_rows((int64_t)reps);
        std::vector<double> D;
        dipole_tensor(r, location, magnetisation, n_particles, ld, D);
        std::vector<double> ma((size_t)rows * ld, 0.0);
        for (size_t rr = 0; rr < reps; ++rr)
            for (size_t p = 0; p < n_particles; ++p)
                for (int c = 0; c < 3; ++c)
                    ma[rr * ld + mp::replica_col((int)p, c)] = m[(rr * n_particles + p) * 3 + c];
        std::vector<double> out((size_t)rows * ld);
        if (mode == MP_DIPOLAR_PAIRS) {
            DevBuf<double> da, db, dc((size_t)rows * ld);
            da.upload(ma);
            db.upload(D);
            const int tot = rows * ld;
            hipLaunchKernelGGL(mp::gemm_f64_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256),
                               0, 0, da.p, db.p, dc.p, rows, ld, ld);
            MP_HIP(hipGetLastError());
            MP_HIP(hipMemcpy(out.data(), dc.p, out.size() * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            DevBuf<float> a32, b32, c32((size_t)rows * ld);
            DevBuf<__bf16> bhi, blo;
            a32.upload(cast_vec<float>(ma));
            if (mode == MP_DIPOLAR_GEMM_F32) {
                b32.upload(cast_vec<float>(D));
            } else {
                std::vector<__bf16> hi, lo;
                split_host(D, hi, lo);
                bhi.upload(hi);
                blo.upload(lo);
            }
            launch_product(mode, a32.p, b32.p, bhi.p, blo.p, c32.p, rows, ld, 0);
            MP_HIP(hipGetLastError());
            std::vector<float> o32(out.size());
            MP_HIP(hipMemcpy(o32.data(), c32.p, o32.size() * sizeof(float), hipMemcpyDeviceToHost));
            out.assign(o32.begin(), o32.end());
        }
        for (size_t rr = 0; rr < reps; ++rr)
            for (size_t p = 0; p < n_particles; ++p)
                for (int c = 0; c < 3; ++c)
                    h[(rr * n_particles + p) * 3 + c] = out[rr * ld + mp::replica_col((int)p, c)];
    });
}

int mp_wiener_normals(uint64_t seed, uint32_t particle, uint32_t tag, uint64_t j0, size_t n,
                      int precision, double *out, int device)
{
    return guarded([&] {
        if (!out) throw simulation::Error(MP_ERR_ARG, "out is required");
        if (n == 0) return;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw simulation::Error(MP_ERR_NODEV, "no HIP device visible");
        if (device < 0 || device >= ndev) throw simulation::Error(MP_ERR_ARG, "bad device");
        MP_HIP(hipSetDevice(device));
        DevBuf<double> d(n);
        const unsigned grid = (unsigned)((n + 255) / 256);
        if (precision == MP_PRECISION_FP32)
            hipLaunchKernelGGL(mp::wiener_kernel<float>, dim3(grid), dim3(256), 0, 0,
                               (uint32_t)seed, (uint32_t)(seed >> 32), particle, tag, j0,
                               (int64_t)n, d.p);
        else
            hipLaunchKernelGGL(mp::wiener_kernel<double>, dim3(grid), dim3(256), 0, 0,
                               (uint32_t)seed, (uint32_t)(seed >> 32), particle, tag, j0,
                               (int64_t)n, d.p);
        MP_HIP(hipGetLastError());
        MP_HIP(hipMemcpy(out, d.p, n * sizeof(double), hipMemcpyDeviceToHost));
    });
}

void mp_constants(double *kb, double *mu0, double *gamma)
{
    if (kb) *kb = kKB;
    if (mu0) *mu0 = kMU0;
    if (gamma) *gamma = kGYROMAG;
}

int mp_save_results(const char *prefix, const double *mx, const double *my, const double *mz,
                    const double *field, const double *time, size_t n_samples)
{
    return guarded([&] {
        require(prefix && mx && my && mz && field && time, "prefix and all five arrays are required");
        const char *ext[5] = {".mx", ".my", ".mz", ".field", ".time"};
        const double *arr[5] = {mx, my, mz, field, time};
        for (int i = 0; i < 5; ++i) {
            const std::string name = std::string(prefix) + ext[i];
            FILE *f = std::fopen(name.c_str(), "wb");
            if (!f) throw simulation::Error(MP_ERR_ARG, "cannot open " + name + " for writing");
            const size_t w = n_samples ? std::fwrite(arr[i], sizeof(double), n_samples, f) : 0;
            const bool ok = std::fclose(f) == 0 && w == n_samples;
            if (!ok) throw simulation::Error(MP_ERR_ARG, "failed to write " + name);
        }
    });
}

const char *mp_last_error(void) { return last_error().c_str(); }

const char *mp_version(void) { return "magpy_amd 0.1 (abi 1, gfx950)"; }

int mp_device_info(int device, char *buf, size_t len)
{
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        last_error() = hipGetErrorString(e);
        return MP_ERR_NODEV;
    }
    if (buf && len)
        std::snprintf(buf, len, "%s (%s), %d CUs, %d MHz, %.1f GiB", prop.name, prop.gcnArchName,
                      prop.multiProcessorCount, prop.clockRate / 1000,
                      (double)prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
    return MP_OK;
}

}  // extern "C"
