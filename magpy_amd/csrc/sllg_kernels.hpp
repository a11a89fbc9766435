// sllg_kernels.hpp -- gfx950 kernels of the stochastic LLG ensemble integrator.
//
// Independent particles (clusters of one, or interactions off): one lane per
// particle, the state and all parameters in registers for a whole step block,
// Wiener increments generated in-lane from the Philox stream (philox.h), and
// samples written SoA ([sample][component][particle]) so a wave's stores are
// contiguous.  Ensemble sums are reduced per wave in FP64 and written as
// per-wave partials, summed afterwards in a fixed order (deterministic).
//
// Physics restated from owlas/magpy:
//   drift     a(m,h) = -m x h + alpha (h |m|^2 - m (m.h))   lib/llg.cpp:14-29
//   diffusion B(m) w = sr [-m x w + alpha (w |m|^2 - m (m.w))] lib/llg.cpp:92-106
//   field     h = k (m.e) e + h_app(t) z                   lib/field.cpp:116-132, 81-88
//   Heun      lib/integrators.cpp:372-405, called with t = (n+1) dt for the
//             step n -> n+1 (lib/simulation.cpp:350-375)
// Both a and B w are the same bilinear map L(m, v) = -m x v + alpha (v |m|^2 -
// m (m.v)) applied to v = h and v = sr w, so one stage costs one L:
//   a dt + B w sqrt(dt) = L(m, h dt + sr sqrt(dt) w).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "linalg3.h"
#include "philox.h"

namespace mp {

constexpr double kPi = 3.14159265358979323846;
constexpr int kBlock = 256;
constexpr int kWave = 64;

enum FieldKind { kSine = 0, kSquare = 1, kConstant = 2, kMixed = 3 };

// ------------------------------------------------------------------ numerics
template <typename T>
struct Num;

template <>
struct Num<float> {
    // Native transcendentals: v_log_f32 (log2), v_sqrt_f32, v_sin/v_cos_f32
    // (argument in revolutions, so sin(2 pi v) is one instruction).
    // -2 ln 2: r^2 = -2 ln u = kRadius2 log2 u
    static constexpr float kRadius2 = -1.38629436111989061883f;
    static __device__ inline void box_muller(uint32_t a, uint32_t b, float &z0, float &z1)
    {
        box_muller_r2(a, b, kRadius2, z0, z1);
    }
    // the pair scaled by s, given c = kRadius2 s^2
    static __device__ inline void box_muller_r2(uint32_t a, uint32_t b, float c, float &z0,
                                                float &z1)
    {
        const float u = fmaf((float)a, 0x1p-32f, 0x1p-33f);
        const float v = fmaf((float)b, 0x1p-32f, 0x1p-33f);
        const float r = __builtin_amdgcn_sqrtf(c * __builtin_amdgcn_logf(u));
        z0 = r * __builtin_amdgcn_cosf(v);
        z1 = r * __builtin_amdgcn_sinf(v);
    }
    // h sin(2 pi c) for c = f t in cycles, reduced to [0,1) in FP64 first.
    static __device__ inline float sine_cycles(double c)
    {
        return __builtin_amdgcn_sinf((float)(c - floor(c)));
    }
};

template <>
struct Num<double> {
    static __device__ inline void box_muller(uint32_t a, uint32_t b, double &z0, double &z1)
    {
        const double u = ((double)a + 0.5) * 0x1p-32;
        const double v = ((double)b + 0.5) * 0x1p-32;
        const double r = sqrt(-2.0 * log(u));
        const double th = 2.0 * kPi * v;
        double s, c;
        sincos(th, &s, &c);
        z0 = r * c;
        z1 = r * s;
    }
};

// Applied field along z at stage time.  FP32 evaluates the phase from cycles
// c = f_red * t in FP64; FP64 evaluates h sin(2 pi f t) with t = (n+1) dt as
// the reference does (lib/field.cpp:37-54).
template <typename T, int kShape>
struct Applied;

template <int kShape>
struct Applied<float, kShape> {
    static __device__ inline float at(int shape, float amp, double cycles)
    {
        const int s = kShape == kMixed ? shape : kShape;
        if (s == kConstant) return amp;
        if (s == kSine) return amp * Num<float>::sine_cycles(cycles);
        return (cycles - floor(cycles)) >= 0.5 ? -amp : amp;  // (int)(2 f t) % 2
    }
};

template <int kShape>
struct Applied<double, kShape> {
    static __device__ inline double at(int shape, double amp, double f, double t)
    {
        const int s = kShape == kMixed ? shape : kShape;
        if (s == kConstant) return amp;
        if (s == kSine) return amp * sin(2 * kPi * f * t);
        return amp * ((int)(t * f * 2) % 2 ? -1 : 1);
    }
};

// a * b + c with one rounding.  The step arithmetic below is written with
// explicit fused multiply-adds under `fp contract(off)`, so which products
// fuse is fixed by the source and not by the surrounding kernel: every kernel
// that evaluates a step (one- and two-wave Heun, the sensitivity forward pass
// and its recompute) produces the same bits from the same inputs.
template <typename T>
__device__ inline T fmad(T a, T b, T c);
template <>
__device__ inline float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <>
__device__ inline double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

// L(m, v) = -m x v + alpha (v |m|^2 - m (m.v)), the reference's drift and
// diffusion applied to v (lib/llg.cpp:14-29, 92-106), evaluated as -w - alpha m x w
// with w = m x v: m x (m x v) = m (m.v) - v |m|^2 for any m, so this is the
// same map in 15 operations (two cross products and three fused
// multiply-adds) instead of 21.
template <typename T>
__device__ inline void llg_map(T mx, T my, T mz, T vx, T vy, T vz, T al, T &lx, T &ly, T &lz)
{
#pragma clang fp contract(off)
    const T wx = fmad(my, vz, -(mz * vy));
    const T wy = fmad(mz, vx, -(mx * vz));
    const T wz = fmad(mx, vy, -(my * vx));
    const T ux = fmad(my, wz, -(mz * wy));
    const T uy = fmad(mz, wx, -(mx * wz));
    const T uz = fmad(mx, wy, -(my * wx));
    lx = fmad(-al, ux, -wx);
    ly = fmad(-al, uy, -wy);
    lz = fmad(-al, uz, -wz);
}

// ------------------------------------------------------------------ arguments
template <typename T>
struct IndepArgs {
    int n;          // particles
    int n_waves;    // waves in the grid (ensemble partial stride)
    // per-particle parameters (member values broadcast to its particles)
    const T *ex, *ey, *ez;  // easy axis
    const T *k;             // K_i / K_avg
    const T *alpha;
    const T *dt;            // reduced step
    const T *sig;           // sr * sqrt(dt)
    const T *sr;            // reduced thermal field strength
    const T *happ;          // reduced field amplitude
    const double *fred;     // reduced field frequency
    const double *fdt;      // f_red * dt_red, cycles per step
    const double *dtd;      // reduced step in FP64 (field time)
    const int *shape;
    const uint32_t *key0, *key1, *pid;
    const double *msw;      // Ms of the particle's member (ensemble sums)
    T *mx, *my, *mz;        // state in/out
    uint32_t *status;
    // sampling
    const int64_t *sched;   // [S] step index recorded as sample s
    int64_t S;
    int64_t s_next;         // first sample with sched[s] >= n0 + 1
    T *samples;             // [S][3][n] or null
    double *ens_partial;    // [S][3][n_waves] or null
    // step block [n0, n1)
    int64_t n0, n1;
    uint32_t tag;
    int renorm;
    const double *noise;    // [steps][n][3] (array noise) or null
};

template <typename T>
__device__ inline void wave_sum3(double &x, double &y, double &z)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        x += __shfl_xor(x, off, kWave);
        y += __shfl_xor(y, off, kWave);
        z += __shfl_xor(z, off, kWave);
    }
}

// wave of the calling lane in a grid of one-wave-per-64-particles blocks
__device__ inline int grid_wave()
{
    return (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
}

// sample s of particle p; `wave` indexes the 64 particles of the calling wave
template <typename T>
__device__ inline void record_sample(const IndepArgs<T> &a, int64_t s, int p, bool live, T mx,
                                     T my, T mz, double w, int wave)
{
    if (a.samples != nullptr && live) {
        T *base = a.samples + (size_t)s * 3 * (size_t)a.n;
        base[p] = mx;
        base[(size_t)a.n + p] = my;
        base[2 * (size_t)a.n + p] = mz;
    }
    if (a.ens_partial != nullptr) {
        double x = live ? (double)mx * w : 0.0;
        double y = live ? (double)my * w : 0.0;
        double z = live ? (double)mz * w : 0.0;
        wave_sum3<T>(x, y, z);
        if ((threadIdx.x & (kWave - 1)) == 0) {
            double *dst = a.ens_partial + (size_t)s * 3 * (size_t)a.n_waves;
            dst[wave] = x;
            dst[(size_t)a.n_waves + wave] = y;
            dst[2 * (size_t)a.n_waves + wave] = z;
        }
    }
}

// ------------------------------------------------------------------ Heun
// One Heun step n -> n+1 of one particle (lib/integrators.cpp:372-405):
// predictor m~ = m + L(m, h(m,t1) dt + s), corrector
// m' = m + (L(m, v1) + L(m~, v2)) / 2 with s = sr sqrt(dt) w shared by both
// stages; h1, h2 are the applied field at the two stage times.  The
// anisotropy field enters as v = (m.ke) edt + s with ke = k e and
// edt = e dt precomputed, which keeps each stage's dependency chain short.
template <typename T, bool kRenorm>
__device__ inline void heun_update(T &mx, T &my, T &mz, T sx, T sy, T sz, T h1, T h2, T kex,
                                   T key, T kez, T edx, T edy, T edz, T dt, T al)
{
#pragma clang fp contract(off)
    const T sz1 = fmad(h1, dt, sz), sz2 = fmad(h2, dt, sz);
    T d = fmad(mx, kex, fmad(my, key, mz * kez));
    T vx = fmad(d, edx, sx), vy = fmad(d, edy, sy), vz = fmad(d, edz, sz1);
    T ax, ay, az;
    llg_map(mx, my, mz, vx, vy, vz, al, ax, ay, az);
    const T tx = mx + ax, ty = my + ay, tz = mz + az;
    d = fmad(tx, kex, fmad(ty, key, tz * kez));
    vx = fmad(d, edx, sx);
    vy = fmad(d, edy, sy);
    vz = fmad(d, edz, sz2);
    T bx, by, bz;
    llg_map(tx, ty, tz, vx, vy, vz, al, bx, by, bz);
    mx = fmad(T(0.5), ax + bx, mx);
    my = fmad(T(0.5), ay + by, my);
    mz = fmad(T(0.5), az + bz, mz);
    if constexpr (kRenorm) {  // lib/simulation.cpp:379-387
        const T nrm = sqrt(fmad(mx, mx, fmad(my, my, mz * mz)));
        mx = mx / nrm;
        my = my / nrm;
        mz = mz / nrm;
    }
}

// Twelve Wiener normals of step group g (steps 4g..4g+3) scaled by sig.
// FP32 folds sig into the Box-Muller radius, sig r = sqrt(-2 ln2 sig^2
// log2 u), which saves two multiplies per pair; with sig = 1 the radius is
// the unscaled one bit for bit.  (sig^2 stays a normal float for any sig
// above 1e-19, far below the noise of any physical temperature.)
template <typename T>
__device__ inline void group_normals(int64_t g, uint32_t pid, uint32_t tag, const PhiloxKeys &keys,
                                     T sig, T z[12])
{
    if constexpr (sizeof(T) == 4) {
        const float c = Num<float>::kRadius2 * (sig * sig);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const uint64_t n = (uint64_t)(3 * g + i);
            const u32x4 w =
                philox4x32_10(u32x4{(uint32_t)n, (uint32_t)(n >> 32), pid, tag}, keys);
            Num<float>::box_muller_r2(w.x, w.y, c, z[4 * i], z[4 * i + 1]);
            Num<float>::box_muller_r2(w.z, w.w, c, z[4 * i + 2], z[4 * i + 3]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const uint64_t n = (uint64_t)(3 * g + i);
            const u32x4 w =
                philox4x32_10(u32x4{(uint32_t)n, (uint32_t)(n >> 32), pid, tag}, keys);
            Num<T>::box_muller(w.x, w.y, z[4 * i], z[4 * i + 1]);
            Num<T>::box_muller(w.z, w.w, z[4 * i + 2], z[4 * i + 3]);
        }
#pragma unroll
        for (int i = 0; i < 12; ++i) z[i] *= sig;
    }
}

// One block of steps [n0, n1) for independent particles.  Steps are walked in
// groups of four, whose twelve normals come from three Philox calls.  The
// state chain of one step depends on the previous step, so with one wave per
// SIMD the kernel is latency bound; the normals of group g+1 are generated in
// the same basic block as the four steps of group g, giving the scheduler an
// independent instruction stream to fill the chain's latency.  Groups that
// are wholly inside the block and record no sample take that fast path.  With
// kTable the applied field of the whole launch is one shared table htab
// (FP32 only): htab[i] = h(t_{4 g0 + 1 + i}) with g0 = n0 / 4 the first
// group, read one group ahead; stage 2 of step n uses h(t_{n+2}), i.e.
// t_{n+1} + dt up to one rounding of t.
template <typename T, int kShape, bool kArrayNoise, bool kTable, bool kRenorm>
__global__ __launch_bounds__(kBlock) void heun_indep_kernel(IndepArgs<T> a,
                                                            const float *__restrict__ htab)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = p < a.n;
    const int q = live ? p : a.n - 1;

    T mx = a.mx[q], my = a.my[q], mz = a.mz[q];
    const T ex = a.ex[q], ey = a.ey[q], ez = a.ez[q];
    const T k = a.k[q], al = a.alpha[q], dt = a.dt[q], sig = a.sig[q], amp = a.happ[q];
    const T kex = k * ex, key = k * ey, kez = k * ez, edx = ex * dt, edy = ey * dt, edz = ez * dt;
    const double fred = a.fred[q], fdt = a.fdt[q], dtd = a.dtd[q];
    const int shape = a.shape[q];
    const double msw = a.msw[q];
    PhiloxKeys keys;
    keys.init(a.key0[q], a.key1[q]);
    const uint32_t pid = a.pid[q];
    const int64_t n0 = a.n0, n1 = a.n1;

    int64_t s = a.s_next;
    int64_t next_rec = s < a.S ? a.sched[s] : INT64_MAX;

    float hc[5], hn[5];  // table entries of the current and next group
    auto fields = [&](int64_t n, int u, double base, T &h1, T &h2) {
        if constexpr (kTable) {
            h1 = (T)hc[u];
            h2 = (T)hc[u + 1];
        } else if constexpr (sizeof(T) == 4) {
            const double c1 = (base + (double)(u + 1)) * fdt;
            h1 = Applied<T, kShape>::at(shape, amp, c1);
            h2 = Applied<T, kShape>::at(shape, amp, c1 + fdt);
        } else {
            const double t1 = (base + (double)(u + 1)) * dtd;
            h1 = Applied<T, kShape>::at(shape, amp, fred, t1);
            h2 = Applied<T, kShape>::at(shape, amp, fred, t1 + dtd);
        }
    };

    T zc[12], zn[12];
    const int64_t g_begin = n0 >> 2, g_end = (n1 + 3) >> 2;
    if (!kArrayNoise) group_normals(g_begin, pid, a.tag, keys, sig, zc);
    if constexpr (kTable) {
#pragma unroll
        for (int j = 0; j < 5; ++j) hc[j] = htab[j];
    }
    for (int64_t g = g_begin; g < g_end; ++g) {
        const double base = (double)(4 * g);
        if (!kArrayNoise) group_normals(g + 1, pid, a.tag, keys, sig, zn);
        if constexpr (kTable) {
            const float *nxt = htab + 4 * (g + 1 - g_begin);
#pragma unroll
            for (int j = 0; j < 5; ++j) hn[j] = nxt[j];
        }
        const bool fast = 4 * g >= n0 && 4 * g + 4 <= n1 && next_rec > 4 * g + 4;
        if (!kArrayNoise && fast) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                T h1, h2;
                fields(4 * g + u, u, base, h1, h2);
                heun_update<T, kRenorm>(mx, my, mz, zc[3 * u], zc[3 * u + 1], zc[3 * u + 2], h1,
                                        h2, kex, key, kez, edx, edy, edz, dt, al);
            }
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t n = 4 * g + u;
                if (n < n0 || n >= n1) continue;
                T sx, sy, sz;
                if (kArrayNoise) {
                    const double *src = a.noise + ((size_t)n * (size_t)a.n + (size_t)q) * 3;
                    sx = sig * (T)src[0];
                    sy = sig * (T)src[1];
                    sz = sig * (T)src[2];
                } else {
                    sx = zc[3 * u];
                    sy = zc[3 * u + 1];
                    sz = zc[3 * u + 2];
                }
                T h1, h2;
                fields(n, u, base, h1, h2);
                heun_update<T, kRenorm>(mx, my, mz, sx, sy, sz, h1, h2, kex, key, kez, edx, edy,
                                        edz, dt, al);
                if (n + 1 == next_rec) {
                    do {
                        record_sample(a, s, p, live, mx, my, mz, msw, grid_wave());
                        ++s;
                        next_rec = s < a.S ? a.sched[s] : INT64_MAX;
                    } while (next_rec == n + 1);
                }
            }
        }
        if (!kArrayNoise) {
#pragma unroll
            for (int i = 0; i < 12; ++i) zc[i] = zn[i];
        }
        if constexpr (kTable) {
#pragma unroll
            for (int j = 0; j < 5; ++j) hc[j] = hn[j];
        }
    }
    if (live) {
        a.mx[p] = mx;
        a.my[p] = my;
        a.mz[p] = mz;
        if (!(isfinite(mx) && isfinite(my) && isfinite(mz))) a.status[p] |= 1u;
    }
}

// ------------------------------------------------------------------ Heun, two waves
// The step block for independent particles with the work of 64
// particles split over a workgroup of two waves: wave 1 (producer) generates
// the noise of 4-step groups — the scaled Wiener normals, or with kBmInt the
// Philox words they are formed from — and, without a shared field table, the
// applied field at both stage times of each step, into an LDS ring of two
// halves of kGroups groups; wave 0 (integrator) runs the Heun steps from the
// other half; the two swap halves at a barrier.  The integrator's dependency
// chain then carries little besides the Heun arithmetic, and independent work
// fills the SIMD cycles the chain leaves idle, which pays where the particles
// are too few to give each SIMD several integrating waves.  Every value is
// computed by the same expression as in heun_indep_kernel, so the results
// are the same bit for bit.
// Ring halves of 4 groups (24 KB, FP32 with a table), 2 groups (20 KB, FP32
// with per-lane fields) or 1 group (20 KB, FP64): four workgroups of a CU fit.
template <typename T, bool kTable>
struct HeunRing {
    static constexpr int kGroups = sizeof(T) == 8 ? 1 : kTable ? 4 : 2;
    static constexpr int kRows = kTable ? 12 : 20;   // normals, then (h1, h2) of each step
};

// htab is deliberately not __restrict__: the integrator then reads it with
// vector loads, which complete in order like its LDS reads, so the reads of
// the next group can be in flight while a group is integrated (scalar loads
// complete out of order and would make every wait a wait for all of them).
// kBmInt (FP32): the producing wave writes the group's Philox words and the
// integrating wave forms the normals (Box-Muller); otherwise the producer
// writes the normals.  With one integrating wave per SIMD the SIMD pairs an
// integrating wave with another workgroup's producing wave, and moving the
// Box-Muller work to the integrating wave fills the stalls of its step chain;
// with fewer integrating waves than SIMDs the integrating wave runs alone and
// the producer should carry it.
template <typename T, int kShape, bool kTable, bool kRenorm, bool kBmInt>
__global__ __launch_bounds__(2 * kWave) void heun_indep_pc_kernel(IndepArgs<T> a,
                                                                  const float *htab)
{
    static_assert(sizeof(T) == 4 || !kTable, "the shared field table is FP32");
    static_assert(sizeof(T) == 4 || !kBmInt, "Box-Muller in the integrating wave is FP32");
    constexpr int K = HeunRing<T, kTable>::kGroups, R = HeunRing<T, kTable>::kRows;
    __shared__ T ring[2][K][R][kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const bool producer = __builtin_amdgcn_readfirstlane(threadIdx.x) >= kWave;
    const int wave = blockIdx.x;
    const int p = wave * kWave + lane;
    const bool live = p < a.n;
    const int q = live ? p : a.n - 1;
    const int64_t n0 = a.n0, n1 = a.n1;
    const int64_t g_begin = n0 >> 2, g_end = (n1 + 3) >> 2;
    const int64_t n_chunks = (g_end - g_begin + K - 1) / K;

    // producer inputs
    const T sig = a.sig[q], amp = a.happ[q];
    const double fdt = a.fdt[q], fred = a.fred[q], dtd = a.dtd[q];
    const int shape = a.shape[q];
    PhiloxKeys keys;
    keys.init(a.key0[q], a.key1[q]);
    const uint32_t pid = a.pid[q];
    // integrator state
    T mx = a.mx[q], my = a.my[q], mz = a.mz[q];
    const T ex = a.ex[q], ey = a.ey[q], ez = a.ez[q];
    const T k = a.k[q], al = a.alpha[q], dt = a.dt[q];
    const T kex = k * ex, key = k * ey, kez = k * ez, edx = ex * dt, edy = ey * dt, edz = ez * dt;
    const double msw = a.msw[q];
    const float cbm = Num<float>::kRadius2 * (float)(sig * sig);   // FP32 Box-Muller radius scale
    int64_t s = a.s_next;
    int64_t next_rec = s < a.S ? a.sched[s] : INT64_MAX;
    // the integrator takes issue priority over co-resident waves
    if (!producer) __builtin_amdgcn_s_setprio(3);
    // pass c: the producer fills chunk c, the integrator runs chunk c - 1
    for (int64_t c = 0; c <= n_chunks; ++c) {
        if (producer) {
            if (c < n_chunks) {
                T (*half)[R][kWave] = ring[c & 1];
                for (int kk = 0; kk < K; ++kk) {
                    const int64_t g = g_begin + c * K + kk;
                    if (g >= g_end) break;
                    if constexpr (kBmInt) {
                        // the group's twelve Philox words (as float bit
                        // patterns); the integrating wave forms the normals
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            const uint64_t cn = (uint64_t)(3 * g + i);
                            const u32x4 w = philox4x32_10(
                                u32x4{(uint32_t)cn, (uint32_t)(cn >> 32), pid, a.tag}, keys);
                            half[kk][4 * i][lane] = __uint_as_float(w.x);
                            half[kk][4 * i + 1][lane] = __uint_as_float(w.y);
                            half[kk][4 * i + 2][lane] = __uint_as_float(w.z);
                            half[kk][4 * i + 3][lane] = __uint_as_float(w.w);
                        }
                    } else {
                        T z[12];
                        group_normals(g, pid, a.tag, keys, sig, z);
#pragma unroll
                        for (int i = 0; i < 12; ++i) half[kk][i][lane] = z[i];
                    }
                    if constexpr (!kTable) {
                        const double base = (double)(4 * g);
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            if constexpr (sizeof(T) == 4) {
                                const double c1 = (base + (double)(u + 1)) * fdt;
                                half[kk][12 + 2 * u][lane] = Applied<T, kShape>::at(shape, amp, c1);
                                half[kk][13 + 2 * u][lane] =
                                    Applied<T, kShape>::at(shape, amp, c1 + fdt);
                            } else {
                                const double t1 = (base + (double)(u + 1)) * dtd;
                                half[kk][12 + 2 * u][lane] =
                                    Applied<T, kShape>::at(shape, amp, fred, t1);
                                half[kk][13 + 2 * u][lane] =
                                    Applied<T, kShape>::at(shape, amp, fred, t1 + dtd);
                            }
                        }
                    }
                }
            }
        } else if (c > 0) {
            T (*half)[R][kWave] = ring[(c - 1) & 1];
            const int64_t gc = g_begin + (c - 1) * K;   // first group of the chunk (< g_end)
            // group kk's ring rows and field values (a group past the end of
            // the block reads the last group's table entries and ring rows it
            // does not use)
            auto load_group = [&](int kk, T *zz, T *hh1, T *hh2) {
#pragma unroll
                for (int i = 0; i < 12; ++i) zz[i] = half[kk][i][lane];
                if constexpr (kTable) {
                    const int64_t gt = gc + kk < g_end ? gc + kk : g_end - 1;
                    const float *tg = htab + 4 * (gt - g_begin);
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        hh1[u] = tg[u];
                        hh2[u] = tg[u + 1];
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        hh1[u] = half[kk][12 + 2 * u][lane];
                        hh2[u] = half[kk][13 + 2 * u][lane];
                    }
                }
            };
            // ring words and field values of every group of the chunk, loaded
            // two groups ahead; the normals of group kk + 1 are formed in the
            // same basic block as the steps of group kk, so their independent
            // work fills the stalls of the step chain.  Unrolled: every index
            // is a constant and the arrays stay in registers.
            T raw[K][12], h1a[K][4], h2a[K][4], zb[K][12];
            auto normals = [&](int kk) {
                if constexpr (kBmInt) {
                    // Box-Muller of the producer's words (group_normals' pairs
                    // and expressions)
#pragma unroll
                    for (int i = 0; i < 6; ++i)
                        Num<float>::box_muller_r2(__float_as_uint(raw[kk][2 * i]),
                                                  __float_as_uint(raw[kk][2 * i + 1]), cbm,
                                                  zb[kk][2 * i], zb[kk][2 * i + 1]);
                } else {
#pragma unroll
                    for (int i = 0; i < 12; ++i) zb[kk][i] = raw[kk][i];
                }
            };
            load_group(0, raw[0], h1a[0], h2a[0]);
            if constexpr (K > 1) load_group(1, raw[1], h1a[1], h2a[1]);
            normals(0);
#pragma unroll
            for (int kk = 0; kk < K; ++kk) {
                const int64_t g = gc + kk;
                if (g >= g_end) break;
                if (kk + 2 < K) load_group(kk + 2, raw[kk + 2], h1a[kk + 2], h2a[kk + 2]);
                const bool fast = 4 * g >= n0 && 4 * g + 4 <= n1 && next_rec > 4 * g + 4;
                if (fast) {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        heun_update<T, kRenorm>(mx, my, mz, zb[kk][3 * u], zb[kk][3 * u + 1],
                                                zb[kk][3 * u + 2], h1a[kk][u], h2a[kk][u], kex, key,
                                                kez, edx, edy, edz, dt, al);
                    if (kk + 1 < K) normals(kk + 1);
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int64_t n = 4 * g + u;
                        if (n < n0 || n >= n1) continue;
                        heun_update<T, kRenorm>(mx, my, mz, zb[kk][3 * u], zb[kk][3 * u + 1],
                                                zb[kk][3 * u + 2], h1a[kk][u], h2a[kk][u], kex,
                                                key, kez, edx, edy, edz, dt, al);
                        if (n + 1 == next_rec) {
                            do {
                                record_sample(a, s, p, live, mx, my, mz, msw, wave);
                                ++s;
                                next_rec = s < a.S ? a.sched[s] : INT64_MAX;
                            } while (next_rec == n + 1);
                        }
                    }
                    if (kk + 1 < K) normals(kk + 1);
                }
            }
        }
        __syncthreads();
    }
    if (producer) return;
    if (live) {
        a.mx[p] = mx;
        a.my[p] = my;
        a.mz[p] = mz;
        if (!(isfinite(mx) && isfinite(my) && isfinite(mz))) a.status[p] |= 1u;
    }
}

// Applied field of a uniform launch: out[i] = h(t_{base+1+i}), i < count,
// evaluated in FP64 from the cycle count c = f_red t reduced to [0,1).
static __global__ void field_table_kernel(int shape, double amp, double fdt, int64_t base, int64_t count,
                                   float *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const double c = (double)(base + 1 + i) * fdt;
    const double fr = c - floor(c);
    double h = amp;
    if (shape == kSine) h = amp * sinpi(2.0 * fr);
    else if (shape == kSquare) h = fr >= 0.5 ? -amp : amp;
    out[i] = (float)h;
}

// ------------------------------------------------------------------ interacting clusters
// Clusters of several particles with dipolar coupling (lib/field.cpp:187-225):
//   h_i += sum_{j != i} w_j / r_ij^3 (3 (m_j.u_ij) u_ij - m_j),  w_j = P v_j,
//   P = mu0 Ms^2 / (8 pi K_avg), u_ij = (p_j - p_i) / |p_j - p_i|,
// with positions already divided by V_avg^(1/3).  Clusters are packed
// first-fit-decreasing into thread blocks, one lane per particle, a cluster
// never straddling two blocks; every lane of a block (padding included) runs
// the same step loop, so block barriers are uniform.  The stage states are
// exchanged through LDS, double-buffered so one barrier per stage suffices.
template <typename T>
struct ClusterArgs {
    IndepArgs<T> base;      // per-slot arrays (n = slots, padding included)
    const T *px, *py, *pz;  // reduced positions
    const T *w;             // P * v_red of the particle
    const int *cbase;       // block-relative lane of the slot's cluster start
    const int *csize;       // cluster size (0 for padding)
};

template <typename T>
struct alignas(16) Vec4 {
    T x, y, z, w;
};

// One dipolar pair term: the field at pi of the moment mj (weight pj.w) at
// pj, added to (ax, ay, az); `skip` masks the particle's own lane.
template <typename T>
__device__ inline void dipolar_term(const Vec4<T> &pi, const Vec4<T> &pj, const Vec4<T> &mj,
                                    bool skip, T &ax, T &ay, T &az)
{
#pragma clang fp contract(off)
    const T dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
    const T r2 = skip ? T(1) : fmad(dx, dx, fmad(dy, dy, dz * dz));
    T rinv;
    if constexpr (sizeof(T) == 4) rinv = __builtin_amdgcn_rsqf(r2);
    else rinv = 1.0 / sqrt(r2);
    const T rinv2 = rinv * rinv;
    const T c = skip ? T(0) : pj.w * rinv2 * rinv;
    const T t = T(3) * fmad(mj.x, dx, fmad(mj.y, dy, mj.z * dz)) * rinv2;
    ax = fmad(c, fmad(t, dx, -mj.x), ax);
    ay = fmad(c, fmad(t, dy, -mj.y), ay);
    az = fmad(c, fmad(t, dz, -mj.z), az);
}

// Dipolar field at lane `self` from lanes [b, e) of its cluster except itself
// (lib/field.cpp:193-202), as two sums taken in increasing j: over the lanes
// at even offsets from b and over those at odd offsets, added at the end.
// With kLanes = 1 one lane carries both sums in one loop; with kLanes = 2 the
// two lanes of a particle (lanes 2i, 2i+1 of the wave) take one sum each
// (`half`) and exchange them.  Either way the result is (even sum) + (odd
// sum), the same bits.  The self term is masked rather than split out, so
// every lane of a cluster runs the same trip count.
template <typename T, int kLanes>
__device__ inline void dipolar_field(const Vec4<T> *pos, const Vec4<T> *m, int self, int b,
                                     int e, int half, T &hx, T &hy, T &hz)
{
    const Vec4<T> pi = pos[self];
    if constexpr (kLanes == 1) {
        T x0 = T(0), y0 = T(0), z0 = T(0), x1 = T(0), y1 = T(0), z1 = T(0);
        int j = b;
        for (; j + 4 <= e; j += 4) {
            Vec4<T> pj[4], mj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                pj[u] = pos[j + u];
                mj[u] = m[j + u];
            }
            dipolar_term(pi, pj[0], mj[0], j == self, x0, y0, z0);
            dipolar_term(pi, pj[1], mj[1], j + 1 == self, x1, y1, z1);
            dipolar_term(pi, pj[2], mj[2], j + 2 == self, x0, y0, z0);
            dipolar_term(pi, pj[3], mj[3], j + 3 == self, x1, y1, z1);
        }
        // j - b is a multiple of 4: the rest start at an even offset
        if (j < e) dipolar_term(pi, pos[j], m[j], j == self, x0, y0, z0);
        if (j + 1 < e) dipolar_term(pi, pos[j + 1], m[j + 1], j + 1 == self, x1, y1, z1);
        if (j + 2 < e) dipolar_term(pi, pos[j + 2], m[j + 2], j + 2 == self, x0, y0, z0);
        hx = x0 + x1;
        hy = y0 + y1;
        hz = z0 + z1;
    } else {
        static_assert(kLanes == 2, "one or two lanes per particle");
        T x = T(0), y = T(0), z = T(0);
        int j = b + half;
        for (; j + 6 < e; j += 8) {
            Vec4<T> pj[4], mj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                pj[u] = pos[j + 2 * u];
                mj[u] = m[j + 2 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) dipolar_term(pi, pj[u], mj[u], j + 2 * u == self, x, y, z);
        }
        for (; j < e; j += 2) dipolar_term(pi, pos[j], m[j], j == self, x, y, z);
        hx = x + __shfl_xor(x, 1);
        hy = y + __shfl_xor(y, 1);
        hz = z + __shfl_xor(z, 1);
    }
}

// The pair geometry of a lane's first K partners with two lanes per particle
// (j = b + half + 2k): positions do not move, so d = p_j - p_i, r^-2 and the
// weight w_j r^-3 are computed once per launch, by the same expressions as
// dipolar_term (the self pair and partners past the cluster's end get weight
// 0), and each stage evaluates only the state-dependent part of those terms.
template <typename T, int K>
struct PairCache {
    T dx[K], dy[K], dz[K], c[K], s[K];
};

template <typename T, int K>
__device__ inline void pair_cache_init(const Vec4<T> *pos, int self, int jb, int e,
                                       PairCache<T, K> &pc)
{
#pragma clang fp contract(off)
    const Vec4<T> pi = pos[self];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = jb + 2 * k;
        const bool in = j < e;
        const Vec4<T> pj = pos[in ? j : self];
        const bool skip = !in || j == self;
        const T dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const T r2 = skip ? T(1) : fmad(dx, dx, fmad(dy, dy, dz * dz));
        const T rinv = __builtin_amdgcn_rsqf(r2);
        const T rinv2 = rinv * rinv;
        pc.dx[k] = skip ? T(0) : dx;
        pc.dy[k] = skip ? T(0) : dy;
        pc.dz[k] = skip ? T(0) : dz;
        pc.s[k] = rinv2;
        pc.c[k] = skip ? T(0) : pj.w * rinv2 * rinv;
    }
}

// The next KL partners' weight w_j r^-3 and r^-2 (j = b + half + 2 (K + k)),
// kept in LDS: geo[k * stride] is this lane's entry k (stride = lanes per
// workgroup), computed by the same expressions as pair_cache_init.
template <typename T, int K, int KL>
__device__ inline void pair_geo_init(const Vec4<T> *pos, int self, int jb, int e, float2 *geo,
                                     int stride)
{
#pragma clang fp contract(off)
    const Vec4<T> pi = pos[self];
    for (int k = 0; k < KL; ++k) {
        const int j = jb + 2 * (K + k);
        const bool in = j < e;
        const Vec4<T> pj = pos[in ? j : self];
        const bool skip = !in || j == self;
        const T dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const T r2 = skip ? T(1) : fmad(dx, dx, fmad(dy, dy, dz * dz));
        const T rinv = __builtin_amdgcn_rsqf(r2);
        const T rinv2 = rinv * rinv;
        geo[k * stride] = float2{skip ? T(0) : pj.w * rinv2 * rinv, rinv2};
    }
}

// Two lanes per particle, the first K partners' geometry cached in
// registers and the next KL partners' weights in LDS (their d = p_j - p_i is
// formed again from the positions): the same terms, in the same order, as
// dipolar_field<T, 2>.
template <typename T, int K, int KL>
__device__ inline void dipolar_field_cached(const Vec4<T> *pos, const Vec4<T> *m, int self, int b,
                                            int e, int half, const PairCache<T, K> &pc,
                                            const float2 *geo, int stride, T &hx, T &hy, T &hz)
{
#pragma clang fp contract(off)
    static_assert(K % 4 == 0, "cached partners come in chunks of four");
    T x = T(0), y = T(0), z = T(0);
    const int jb = b + half;
#pragma unroll
    for (int c4 = 0; c4 < K / 4; ++c4) {
        if (jb + 8 * c4 < e) {
            Vec4<T> mj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = jb + 2 * (4 * c4 + u);
                mj[u] = m[j < e ? j : self];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = 4 * c4 + u;
                const T t = T(3) * fmad(mj[u].x, pc.dx[k], fmad(mj[u].y, pc.dy[k], mj[u].z * pc.dz[k])) *
                            pc.s[k];
                x = fmad(pc.c[k], fmad(t, pc.dx[k], -mj[u].x), x);
                y = fmad(pc.c[k], fmad(t, pc.dy[k], -mj[u].y), y);
                z = fmad(pc.c[k], fmad(t, pc.dz[k], -mj[u].z), z);
            }
        }
    }
    const Vec4<T> pi = pos[self];
    static_assert(KL % 2 == 0, "LDS partners come in pairs");
#pragma unroll
    for (int c2 = 0; c2 < KL / 2; ++c2) {
        if (jb + 2 * (K + 2 * c2) < e) {
            Vec4<T> pj[2], mj[2];
            float2 g[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int j = jb + 2 * (K + 2 * c2 + u);
                const int js = j < e ? j : self;
                pj[u] = pos[js];
                mj[u] = m[js];
                g[u] = geo[(2 * c2 + u) * stride];
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const T dx = pj[u].x - pi.x, dy = pj[u].y - pi.y, dz = pj[u].z - pi.z;
                const T t = T(3) * fmad(mj[u].x, dx, fmad(mj[u].y, dy, mj[u].z * dz)) * g[u].y;
                x = fmad(g[u].x, fmad(t, dx, -mj[u].x), x);
                y = fmad(g[u].x, fmad(t, dy, -mj[u].y), y);
                z = fmad(g[u].x, fmad(t, dz, -mj[u].z), z);
            }
        }
    }
    int j = jb + 2 * (K + KL);
    for (; j + 6 < e; j += 8) {
        Vec4<T> pj[4], mj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            pj[u] = pos[j + 2 * u];
            mj[u] = m[j + 2 * u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) dipolar_term(pi, pj[u], mj[u], j + 2 * u == self, x, y, z);
    }
    for (; j < e; j += 2) dipolar_term(pi, pos[j], m[j], j == self, x, y, z);
    hx = x + __shfl_xor(x, 1);
    hy = y + __shfl_xor(y, 1);
    hz = z + __shfl_xor(z, 1);
}

// kLanes lanes per slot: the block holds blockDim.x / kLanes slots; with two,
// the lanes of a slot split its pair sums and repeat the rest of the step, and
// the even lane alone writes the shared state, samples and results.  With
// kArrayNoise the Wiener normals are the caller's ([step][slot][3], the
// reference's RngArray, lib/rng.cpp:67-94) instead of the Philox stream.
// kCache (two lanes, FP32) caches the first kCache partners' geometry per
// lane (PairCache): workgroups of at most 256 lanes then, for the registers;
// kLds more partners keep their weights in LDS after the three state arrays
// (kLds x lanes float2).
template <typename T, int kShape, bool kTable, bool kRenorm, int kLanes, bool kArrayNoise,
          int kCache = 0, int kLds = 0>
__global__ __launch_bounds__(kCache > 0 ? 256 : 1024, kLds > 0 ? 2 : 1) void heun_cluster_kernel(
    ClusterArgs<T> ca, const float *__restrict__ htab)
{
#pragma clang fp contract(off)
    const IndepArgs<T> &a = ca.base;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int slots = blockDim.x / kLanes;
    Vec4<T> *pos = reinterpret_cast<Vec4<T> *>(smem_raw);
    Vec4<T> *mb0 = pos + slots;
    Vec4<T> *mb1 = mb0 + slots;

    const int tid = threadIdx.x / kLanes;           // block-relative slot
    const int half = threadIdx.x % kLanes;
    const bool lead = half == 0;
    const int p = blockIdx.x * slots + tid;   // slot: the grid covers all slots exactly
    const int cn = ca.csize[p];
    const bool live = cn > 0;
    const int cb = ca.cbase[p], ce = cb + cn;

    T mx = a.mx[p], my = a.my[p], mz = a.mz[p];
    const T ex = a.ex[p], ey = a.ey[p], ez = a.ez[p];
    const T k = a.k[p], al = a.alpha[p], dt = a.dt[p], sig = a.sig[p], amp = a.happ[p];
    const T kex = k * ex, key = k * ey, kez = k * ez, edx = ex * dt, edy = ey * dt, edz = ez * dt;
    const double fred = a.fred[p], fdt = a.fdt[p], dtd = a.dtd[p];
    const int shape = a.shape[p];
    const double msw = a.msw[p];
    PhiloxKeys keys;
    keys.init(a.key0[p], a.key1[p]);
    const uint32_t pid = a.pid[p];
    const int64_t n0 = a.n0, n1 = a.n1;
    const int64_t tab0 = 4 * (n0 >> 2);
    if (lead) pos[tid] = Vec4<T>{ca.px[p], ca.py[p], ca.pz[p], ca.w[p]};
    static_assert(kCache == 0 || (kLanes == 2 && sizeof(T) == 4), "cache: two lanes, FP32");
    static_assert(kLds == 0 || kCache > 0, "LDS partners follow the register-cached ones");
    PairCache<T, (kCache > 0 ? kCache : 4)> pc;
    float2 *geo = reinterpret_cast<float2 *>(mb1 + slots) + threadIdx.x;
    if constexpr (kCache > 0) {
        __syncthreads();
        pair_cache_init(pos, tid, cb + half, ce, pc);
        if constexpr (kLds > 0) pair_geo_init<T, kCache, kLds>(pos, tid, cb + half, ce, geo, blockDim.x);
    }

    int64_t s = a.s_next;
    int64_t next_rec = s < a.S ? a.sched[s] : INT64_MAX;

    // the normals of step group g+1 are drawn while group g is integrated
    T zc[12], zn[12];
    const int64_t g_begin = n0 >> 2, g_end = (n1 + 3) >> 2;
    if constexpr (!kArrayNoise) group_normals(g_begin, pid, a.tag, keys, sig, zc);
    for (int64_t g = g_begin; g < g_end; ++g) {
        if constexpr (!kArrayNoise) group_normals(g + 1, pid, a.tag, keys, sig, zn);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t n = 4 * g + u;
            if (n < n0 || n >= n1) continue;
            T sx, sy, sz;
            if constexpr (kArrayNoise) {
                const double *src = a.noise + ((size_t)n * (size_t)a.n + (size_t)p) * 3;
                sx = sig * (T)src[0];
                sy = sig * (T)src[1];
                sz = sig * (T)src[2];
            } else {
                sx = zc[3 * u];
                sy = zc[3 * u + 1];
                sz = zc[3 * u + 2];
            }
            T h1, h2;
            if constexpr (kTable) {
                h1 = (T)htab[n - tab0];
                h2 = (T)htab[n - tab0 + 1];
            } else if constexpr (sizeof(T) == 4) {
                const double c1 = (double)(n + 1) * fdt;
                h1 = Applied<T, kShape>::at(shape, amp, c1);
                h2 = Applied<T, kShape>::at(shape, amp, c1 + fdt);
            } else {
                const double t1 = (double)(n + 1) * dtd;
                h1 = Applied<T, kShape>::at(shape, amp, fred, t1);
                h2 = Applied<T, kShape>::at(shape, amp, fred, t1 + dtd);
            }
            // stage 1
            if (lead) mb0[tid] = Vec4<T>{mx, my, mz, T(0)};
            __syncthreads();
            T hx, hy, hz;
            if constexpr (kCache > 0)
                dipolar_field_cached<T, kCache, kLds>(pos, mb0, tid, cb, ce, half, pc, geo,
                                                      blockDim.x, hx, hy, hz);
            else
                dipolar_field<T, kLanes>(pos, mb0, tid, cb, ce, half, hx, hy, hz);
            T d = fmad(mx, kex, fmad(my, key, mz * kez));
            T vx = fmad(d, edx, fmad(hx, dt, sx)), vy = fmad(d, edy, fmad(hy, dt, sy)),
              vz = fmad(d, edz, fmad(hz + h1, dt, sz));
            T ax, ay, az;
            llg_map(mx, my, mz, vx, vy, vz, al, ax, ay, az);
            const T tx = mx + ax, ty = my + ay, tz = mz + az;
            // stage 2
            if (lead) mb1[tid] = Vec4<T>{tx, ty, tz, T(0)};
            __syncthreads();
            if constexpr (kCache > 0)
                dipolar_field_cached<T, kCache, kLds>(pos, mb1, tid, cb, ce, half, pc, geo,
                                                      blockDim.x, hx, hy, hz);
            else
                dipolar_field<T, kLanes>(pos, mb1, tid, cb, ce, half, hx, hy, hz);
            d = fmad(tx, kex, fmad(ty, key, tz * kez));
            vx = fmad(d, edx, fmad(hx, dt, sx));
            vy = fmad(d, edy, fmad(hy, dt, sy));
            vz = fmad(d, edz, fmad(hz + h2, dt, sz));
            T bx, by, bz;
            llg_map(tx, ty, tz, vx, vy, vz, al, bx, by, bz);
            mx = fmad(T(0.5), ax + bx, mx);
            my = fmad(T(0.5), ay + by, my);
            mz = fmad(T(0.5), az + bz, mz);
            if constexpr (kRenorm) {
                if (live) {
                    const T nrm = sqrt(fmad(mx, mx, fmad(my, my, mz * mz)));
                    mx = mx / nrm;
                    my = my / nrm;
                    mz = mz / nrm;
                }
            }
            if (n + 1 == next_rec) {
                do {
                    record_sample(a, s, p, live && lead, mx, my, mz, msw, grid_wave());
                    ++s;
                    next_rec = s < a.S ? a.sched[s] : INT64_MAX;
                } while (next_rec == n + 1);
            }
        }
        if constexpr (!kArrayNoise) {
#pragma unroll
            for (int i = 0; i < 12; ++i) zc[i] = zn[i];
        }
    }
    if (!lead) return;
    a.mx[p] = mx;
    a.my[p] = my;
    a.mz[p] = mz;
    if (live && !(isfinite(mx) && isfinite(my) && isfinite(mz))) a.status[p] |= 1u;
}

// ------------------------------------------------------------------ implicit midpoint
// FP64, one lane per particle.  Restates integrator::implicit_midpoint
// (lib/integrators.cpp:576-651) with optimisation::newton_raphson_noinv
// (lib/optimisation.cpp:81-149) and the LLG Jacobians of lib/llg.cpp:68-81,
// 118-158, including the reference's Jacobian as it is written (drift term
// without dt; two diffusion-Jacobian entries taken at m2): they steer the
// Newton path, not the root it converges to.
struct ImplicitArgs {
    IndepArgs<double> base;
    double eps;
    int max_iter;
    unsigned long long *newton_iters;   // per-lane accumulators [n]
    unsigned long long *newton_fail;    // [n]
};

__device__ inline void drift3(const double *m, double al, const double *h, double *d)
{
    d[0] = m[2] * h[1] - m[1] * h[2] + al * (h[0] * (m[1] * m[1] + m[2] * m[2]) - m[0] * (m[1] * h[1] + m[2] * h[2]));
    d[1] = m[0] * h[2] - m[2] * h[0] + al * (h[1] * (m[0] * m[0] + m[2] * m[2]) - m[1] * (m[0] * h[0] + m[2] * h[2]));
    d[2] = m[1] * h[0] - m[0] * h[1] + al * (h[2] * (m[0] * m[0] + m[1] * m[1]) - m[2] * (m[0] * h[0] + m[1] * h[1]));
}

__device__ inline void diffusion3(const double *m, double sr, double al, double *B)
{
    B[0] = al * sr * (m[1] * m[1] + m[2] * m[2]);
    B[1] = sr * (m[2] - al * m[0] * m[1]);
    B[2] = -sr * (m[1] + al * m[0] * m[2]);
    B[3] = -sr * (m[2] + al * m[0] * m[1]);
    B[4] = al * sr * (m[0] * m[0] + m[2] * m[2]);
    B[5] = sr * (m[0] - al * m[1] * m[2]);
    B[6] = sr * (m[1] - al * m[0] * m[2]);
    B[7] = -sr * (m[0] + al * m[1] * m[2]);
    B[8] = al * sr * (m[0] * m[0] + m[1] * m[1]);
}

// d(drift)/dm, row-major, hj = field Jacobian (row-major)
__device__ inline void drift_jac3(const double *m, double a, const double *h, const double *hj, double *j)
{
    j[0] = m[2] * hj[3] - m[1] * hj[6] + a * (-m[1] * h[1] - m[2] * h[2] + (m[1] * m[1] + m[2] * m[2]) * hj[0] - m[0] * (m[1] * hj[3] + m[2] * hj[6]));
    j[1] = -h[2] + m[2] * hj[4] - m[1] * hj[7] + a * (2 * m[1] * h[0] + (m[1] * m[1] + m[2] * m[2]) * hj[1] - m[0] * (h[1] + m[1] * hj[4] + m[2] * hj[7]));
    j[2] = h[1] + m[2] * hj[5] - m[1] * hj[8] + a * (2 * m[2] * h[0] + (m[1] * m[1] + m[2] * m[2]) * hj[2] - m[0] * (h[2] + m[1] * hj[5] + m[2] * hj[8]));
    j[3] = h[2] - m[2] * hj[0] + m[0] * hj[6] + a * (2 * m[0] * h[1] + (m[0] * m[0] + m[2] * m[2]) * hj[3] - m[1] * (h[0] + m[0] * hj[0] + m[2] * hj[6]));
    j[4] = -m[2] * hj[1] + m[0] * hj[7] + a * (-m[0] * h[0] - m[2] * h[2] + (m[0] * m[0] + m[2] * m[2]) * hj[4] - m[1] * (m[0] * hj[1] + m[2] * hj[7]));
    j[5] = -h[0] - m[2] * hj[2] + m[0] * hj[8] + a * (2 * m[2] * h[1] + (m[0] * m[0] + m[2] * m[2]) * hj[5] - m[1] * (h[2] + m[0] * hj[2] + m[2] * hj[8]));
    j[6] = -h[1] + m[1] * hj[0] - m[0] * hj[3] + a * (2 * m[0] * h[2] + (m[0] * m[0] + m[1] * m[1]) * hj[6] - m[2] * (h[0] + m[0] * hj[0] + m[1] * hj[3]));
    j[7] = h[0] + m[1] * hj[1] - m[0] * hj[4] + a * (2 * m[1] * h[2] + (m[0] * m[0] + m[1] * m[1]) * hj[7] - m[2] * (h[1] + m[0] * hj[1] + m[1] * hj[4]));
    j[8] = m[1] * hj[2] - m[0] * hj[5] + a * (-m[0] * h[0] - m[1] * h[1] + (m[0] * m[0] + m[1] * m[1]) * hj[8] - m[2] * (m[0] * hj[2] + m[1] * hj[5]));
}

// jb[z + 3y + 9x] = dB[x][y]/dm[z] as lib/llg.cpp:118-158 writes it
__device__ inline void diffusion_jac3(const double *m, double sr, double al, double *jb)
{
    jb[0] = 0;                  jb[1] = 2 * al * sr * m[1];  jb[2] = 2 * al * sr * m[2];
    jb[3] = -al * sr * m[1];    jb[4] = -al * sr * m[2];     jb[5] = sr;
    jb[6] = -al * sr * m[2];    jb[7] = -sr;                 jb[8] = -al * sr * m[0];
    jb[9] = -al * sr * m[1];    jb[10] = -al * sr * m[0];    jb[11] = -sr;
    jb[12] = 2 * al * sr * m[0]; jb[13] = 0;                 jb[14] = 2 * al * sr * m[2];
    jb[15] = sr;                jb[16] = -al * sr * m[2];    jb[17] = -al * sr * m[1];
    jb[18] = -al * sr * m[2];   jb[19] = sr;                 jb[20] = -al * sr * m[0];
    jb[21] = -sr;               jb[22] = -al * sr * m[2];    jb[23] = -al * sr * m[1];
    jb[24] = 2 * al * sr * m[0]; jb[25] = 2 * al * sr * m[2]; jb[26] = 0;
}

template <int kShape, bool kArrayNoise>
__global__ __launch_bounds__(kBlock) void implicit_indep_kernel(ImplicitArgs ia)
{
    const IndepArgs<double> &a = ia.base;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = p < a.n;
    const int q = live ? p : a.n - 1;

    double x0[3] = {a.mx[q], a.my[q], a.mz[q]};
    const double e[3] = {a.ex[q], a.ey[q], a.ez[q]};
    const double k = a.k[q], al = a.alpha[q], dt = a.dt[q], amp = a.happ[q];
    const double sr = a.sr[q];
    const double fred = a.fred[q];
    const int shape = a.shape[q];
    const double msw = a.msw[q];
    PhiloxKeys keys;
    keys.init(a.key0[q], a.key1[q]);
    const uint32_t pid = a.pid[q];
    const double ah = sqrt(2 * 1000.0 * fabs(log(dt)));
    const double sdt = sqrt(dt);
    // field Jacobian k e e^T (lib/field.cpp:159-174)
    double hj[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) hj[i * 3 + j] = k * e[i] * e[j];

    unsigned long long iters = 0, fails = 0;
    int64_t s = a.s_next;
    int64_t next_rec = s < a.S ? a.sched[s] : INT64_MAX;
    double zbuf[4];
    uint64_t have = UINT64_MAX;

    for (int64_t n = a.n0; n < a.n1; ++n) {
        double w[3];
        if (kArrayNoise) {
            const double *src = a.noise + ((size_t)n * (size_t)a.n + (size_t)q) * 3;
            w[0] = src[0]; w[1] = src[1]; w[2] = src[2];
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const uint64_t j = 3 * (uint64_t)n + d;
                const uint64_t c = j >> 2;
                if (c != have) {
                    const u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)(c >> 32), pid, a.tag}, keys);
                    Num<double>::box_muller(r.x, r.y, zbuf[0], zbuf[1]);
                    Num<double>::box_muller(r.z, r.w, zbuf[2], zbuf[3]);
                    have = c;
                }
                const int jj = (int)(j & 3);
                w[d] = jj == 0 ? zbuf[0] : jj == 1 ? zbuf[1] : jj == 2 ? zbuf[2] : zbuf[3];
            }
        }
        const double t = (double)(n + 1) * dt;
        double dwm[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) dwm[d] = fmax(-ah, fmin(ah, w[d])) * sdt;

        // Euler guess from sde(x0, t) (lib/integrators.cpp:605-614)
        double h[3], dr[3], B[9];
        {
            const double dot = (x0[0] * e[0] + x0[1] * e[1] + x0[2] * e[2]) * k;
            h[0] = dot * e[0];
            h[1] = dot * e[1];
            h[2] = dot * e[2] + Applied<double, kShape>::at(shape, amp, fred, t);
        }
        drift3(x0, al, h, dr);
        diffusion3(x0, sr, al, B);
        double y[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double g = dr[i] * dt;
#pragma unroll
            for (int j = 0; j < 3; ++j) g += B[i * 3 + j] * dwm[j];
            y[i] = (g + x0[i]) / 2;
        }
        const double happ_mid = Applied<double, kShape>::at(shape, amp, fred, t + dt / 2);
        const double tol = ia.eps * sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        double err = 2 * tol;
        int iter = ia.max_iter;
        bool singular = false;
        while ((err > tol) && (iter-- > 0)) {
            const double xt[3] = {y[0], y[1], y[2]};
            const double dot = (xt[0] * e[0] + xt[1] * e[1] + xt[2] * e[2]) * k;
            h[0] = dot * e[0];
            h[1] = dot * e[1];
            h[2] = dot * e[2] + happ_mid;
            drift3(xt, al, h, dr);
            diffusion3(xt, sr, al, B);
            double J[9], jb[27], F[3];
            drift_jac3(xt, al, h, hj, J);
            diffusion_jac3(xt, sr, al, jb);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                F[i] = xt[i] - 0.5 * dr[i] * dt - x0[i];
#pragma unroll
                for (int j = 0; j < 3; ++j) F[i] -= 0.5 * B[i * 3 + j] * dwm[j];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double v = (double)(i == j) - 0.5 * J[i * 3 + j];
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk) v -= 0.5 * jb[i * 9 + kk * 3 + j] * dwm[kk];
                    J[i * 3 + j] = v;
                }
            double dlt[3] = {-F[0], -F[1], -F[2]};
            ++iters;
            if (!solve3(J, dlt)) {  // LAPACK_ERR: x_root left at -F
                y[0] = dlt[0]; y[1] = dlt[1]; y[2] = dlt[2];
                singular = true;
                break;
            }
            err = sqrt(dlt[0] * dlt[0] + dlt[1] * dlt[1] + dlt[2] * dlt[2]);
            y[0] = dlt[0] + xt[0];
            y[1] = dlt[1] + xt[1];
            y[2] = dlt[2] + xt[2];
        }
        if (singular || iter == -1) ++fails;
#pragma unroll
        for (int i = 0; i < 3; ++i) x0[i] = 2 * y[i] - x0[i];
        if (a.renorm) {
            const double nrm = sqrt(x0[0] * x0[0] + x0[1] * x0[1] + x0[2] * x0[2]);
            x0[0] = x0[0] / nrm; x0[1] = x0[1] / nrm; x0[2] = x0[2] / nrm;
        }
        if (n + 1 == next_rec) {
            do {
                record_sample(a, s, p, live, x0[0], x0[1], x0[2], msw, grid_wave());
                ++s;
                next_rec = s < a.S ? a.sched[s] : INT64_MAX;
            } while (next_rec == n + 1);
        }
    }
    if (live) {
        a.mx[p] = x0[0];
        a.my[p] = x0[1];
        a.mz[p] = x0[2];
        ia.newton_iters[p] += iters;
        ia.newton_fail[p] += fails;
        uint32_t st = 0;
        if (!(isfinite(x0[0]) && isfinite(x0[1]) && isfinite(x0[2]))) st |= 1u;
        if (fails) st |= 2u;
        a.status[p] |= st;
    }
}

// ------------------------------------------------------------------ implicit midpoint, clusters
// Interacting clusters in FP64, one lane per particle, whole clusters per
// workgroup (packed as for the Heun pair sums).  The reference takes one
// Newton step of a 3N x 3N system per cluster step (lib/integrators.cpp:
// 576-651) whose Jacobian is block diagonal: the drift Jacobian keeps each
// particle's own block only (llg::multi_drift_quasijacobian,
// lib/llg.cpp:378-402; the field Jacobian leaves out interactions,
// lib/simulation.cpp:292-303) and the diffusion Jacobian is per particle
// (lib/llg.cpp:409-427).  LU with partial pivoting of a block-diagonal
// matrix never pivots across blocks, and its eliminations and substitutions
// only subtract exact zeros outside a block, so the 3N solve is the N 3x3
// solves of its blocks (solve3) value for value.  What couples the
// particles is F: the effective field at every iterate includes the dipolar
// sum over the cluster's iterates, exchanged through LDS each iteration; and
// the tolerance eps |guess|_2 and the step norm |delta|_2 are taken over the
// whole cluster (lib/optimisation.cpp:98, 142), summed over its members in
// component order, so the lanes of a cluster iterate together.  A singular
// block makes the whole system singular: dgesv then leaves its right-hand
// side -F as the result for every particle of the cluster.  Each particle's
// drift Jacobian reads the field Jacobian at hj + 3n, n its index in the
// cluster (lib/llg.cpp:387): a slice of row 0 (and row 1) of the 3N x 3N
// block-diagonal k e e^T, which for n >= 1 is not the particle's own block.
// It is kept as written: it steers Newton's path and iteration count, not
// the root.  Newton counts are kept on each cluster's first lane, so their
// sum over lanes is the sum over clusters of the reference's counts.  Without
// interactions the dipolar sums are left out and the cluster still shares
// its norms, as the reference's 3N system does.
struct ImplicitClusterArgs {
    ClusterArgs<double> c;
    double eps;
    int max_iter;
    unsigned long long *newton_iters;   // [slots]
    unsigned long long *newton_fail;    // [slots]
    int interactions;                   // 0: no dipolar sums (norms stay per cluster)
};

constexpr int kImplicitClusterBlock = 256;   // largest workgroup, so largest cluster

template <int kShape, bool kArrayNoise>
__global__ __launch_bounds__(kImplicitClusterBlock) void implicit_cluster_kernel(
    ImplicitClusterArgs ia)
{
    const ClusterArgs<double> &ca = ia.c;
    const IndepArgs<double> &a = ca.base;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int slots = blockDim.x;
    Vec4<double> *pos = reinterpret_cast<Vec4<double> *>(smem_raw);
    Vec4<double> *mb = pos + slots;    // the states the field is taken from
    Vec4<double> *red = mb + slots;    // vectors whose cluster norm is taken

    const int tid = threadIdx.x;
    const int p = blockIdx.x * slots + tid;   // the grid covers all slots exactly
    const int cn = ca.csize[p];
    const bool live = cn > 0;
    const int cb = live ? ca.cbase[p] : tid, ce = live ? cb + cn : tid;
    const bool leader = live && tid == cb;

    double x0[3] = {a.mx[p], a.my[p], a.mz[p]};
    const double e[3] = {a.ex[p], a.ey[p], a.ez[p]};
    const double k = a.k[p], al = a.alpha[p], dt = a.dt[p], amp = a.happ[p];
    const double sr = a.sr[p], fred = a.fred[p];
    const int shape = a.shape[p];
    const double msw = a.msw[p];
    PhiloxKeys keys;
    keys.init(a.key0[p], a.key1[p]);
    const uint32_t pid = a.pid[p];
    const double ah = sqrt(2 * 1000.0 * fabs(log(dt)));
    const double sdt = sqrt(dt);
    pos[tid] = Vec4<double>{ca.px[p], ca.py[p], ca.pz[p], ca.w[p]};

    // hj + 3n of the 3N x 3N field Jacobian J[i][j] = k_{i/3} e[i] e[j] for
    // i/3 == j/3 (lib/field.cpp:159-174): flat index f = 3n + q, row f / 3N,
    // column f % 3N.  For every n < N that row is 0, 1 or 2, i.e. particle 0's
    // block row, nonzero in columns 0-2 only.
    double hj[9];
    {
        const int nn = live ? cn : 1, idx = live ? tid - cb : 0;
        const int first = blockIdx.x * slots + cb;
        const double k0 = a.k[first];
        const double e0[3] = {a.ex[first], a.ey[first], a.ez[first]};
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int f = 3 * idx + q, i = f / (3 * nn), j = f % (3 * nn);
            hj[q] = (i < 3 && j < 3) ? k0 * e0[i < 3 ? i : 0] * e0[j < 3 ? j : 0] : 0.0;
        }
    }

    unsigned long long iters = 0, fails = 0;
    int64_t s = a.s_next;
    int64_t next_rec = s < a.S ? a.sched[s] : INT64_MAX;
    double zbuf[4];
    uint64_t have = UINT64_MAX;

    for (int64_t n = a.n0; n < a.n1; ++n) {
        double w[3];
        if (kArrayNoise) {
            const double *src = a.noise + ((size_t)n * (size_t)a.n + (size_t)p) * 3;
            w[0] = src[0]; w[1] = src[1]; w[2] = src[2];
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const uint64_t j = 3 * (uint64_t)n + d;
                const uint64_t c = j >> 2;
                if (c != have) {
                    const u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)(c >> 32), pid, a.tag}, keys);
                    Num<double>::box_muller(r.x, r.y, zbuf[0], zbuf[1]);
                    Num<double>::box_muller(r.z, r.w, zbuf[2], zbuf[3]);
                    have = c;
                }
                const int jj = (int)(j & 3);
                w[d] = jj == 0 ? zbuf[0] : jj == 1 ? zbuf[1] : jj == 2 ? zbuf[2] : zbuf[3];
            }
        }
        const double t = (double)(n + 1) * dt;
        double dwm[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) dwm[d] = fmax(-ah, fmin(ah, w[d])) * sdt;

        // Euler guess from sde(x0, t), the field including the cluster's
        // dipolar sum at x0 (lib/integrators.cpp:605-614)
        mb[tid] = Vec4<double>{x0[0], x0[1], x0[2], 0.0};
        __syncthreads();
        double h[3], hd[3] = {0.0, 0.0, 0.0}, dr[3], B[9], y[3];
        if (ia.interactions) dipolar_field<double, 1>(pos, mb, tid, cb, ce, 0, hd[0], hd[1], hd[2]);
        {
            const double dot = (x0[0] * e[0] + x0[1] * e[1] + x0[2] * e[2]) * k;
            h[0] = dot * e[0];
            h[1] = dot * e[1];
            h[2] = dot * e[2] + Applied<double, kShape>::at(shape, amp, fred, t);
            if (ia.interactions) {
                h[0] = h[0] + hd[0];
                h[1] = h[1] + hd[1];
                h[2] = h[2] + hd[2];
            }
        }
        drift3(x0, al, h, dr);
        diffusion3(x0, sr, al, B);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double g = dr[i] * dt;
#pragma unroll
            for (int j = 0; j < 3; ++j) g += B[i * 3 + j] * dwm[j];
            y[i] = (g + x0[i]) / 2;
        }
        red[tid] = Vec4<double>{y[0], y[1], y[2], 0.0};
        __syncthreads();   // every lane's guess field is taken: mb may change
        double ss = 0.0;
        for (int j = cb; j < ce; ++j) {
            const Vec4<double> v = red[j];
            ss += v.x * v.x;
            ss += v.y * v.y;
            ss += v.z * v.z;
        }
        const double tol = ia.eps * sqrt(ss);
        mb[tid] = Vec4<double>{y[0], y[1], y[2], 0.0};
        const double happ_mid = Applied<double, kShape>::at(shape, amp, fred, t + dt / 2);
        double err = 2 * tol;
        int iter = ia.max_iter;
        bool singular = false;
        bool go = live && (err > tol) && (iter-- > 0);
        while (__syncthreads_or(go)) {
            const double xt[3] = {y[0], y[1], y[2]};
            double dlt[3] = {0.0, 0.0, 0.0}, nf[3] = {0.0, 0.0, 0.0};
            bool ok = true;
            if (go) {
                const double dot = (xt[0] * e[0] + xt[1] * e[1] + xt[2] * e[2]) * k;
                h[0] = dot * e[0];
                h[1] = dot * e[1];
                h[2] = dot * e[2] + happ_mid;
                if (ia.interactions) {
                    dipolar_field<double, 1>(pos, mb, tid, cb, ce, 0, hd[0], hd[1], hd[2]);
                    h[0] = h[0] + hd[0];
                    h[1] = h[1] + hd[1];
                    h[2] = h[2] + hd[2];
                }
                drift3(xt, al, h, dr);
                diffusion3(xt, sr, al, B);
                double J[9], jb[27], F[3];
                drift_jac3(xt, al, h, hj, J);
                diffusion_jac3(xt, sr, al, jb);
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    F[i] = xt[i] - 0.5 * dr[i] * dt - x0[i];
#pragma unroll
                    for (int j = 0; j < 3; ++j) F[i] -= 0.5 * B[i * 3 + j] * dwm[j];
                }
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        double v = (double)(i == j) - 0.5 * J[i * 3 + j];
#pragma unroll
                        for (int kk = 0; kk < 3; ++kk) v -= 0.5 * jb[i * 9 + kk * 3 + j] * dwm[kk];
                        J[i * 3 + j] = v;
                    }
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    dlt[i] = -F[i];
                    nf[i] = -F[i];
                }
                ++iters;
                ok = solve3(J, dlt);
            }
            __syncthreads();   // the iterates' field and the last norm are read
            if (go) red[tid] = Vec4<double>{dlt[0], dlt[1], dlt[2], ok ? 0.0 : 1.0};
            __syncthreads();
            if (go) {
                double s2 = 0.0;
                bool sing = false;
                for (int j = cb; j < ce; ++j) {
                    const Vec4<double> v = red[j];
                    s2 += v.x * v.x;
                    s2 += v.y * v.y;
                    s2 += v.z * v.z;
                    sing = sing || v.w != 0.0;
                }
                if (sing) {   // LAPACK_ERR: x_root left at -F
                    y[0] = nf[0]; y[1] = nf[1]; y[2] = nf[2];
                    singular = true;
                    go = false;
                } else {
                    err = sqrt(s2);
                    y[0] = dlt[0] + xt[0];
                    y[1] = dlt[1] + xt[1];
                    y[2] = dlt[2] + xt[2];
                    mb[tid] = Vec4<double>{y[0], y[1], y[2], 0.0};
                    go = (err > tol) && (iter-- > 0);
                }
            }
        }
        if (live && (singular || iter == -1)) ++fails;
#pragma unroll
        for (int i = 0; i < 3; ++i) x0[i] = 2 * y[i] - x0[i];
        if (a.renorm && live) {
            const double nrm = sqrt(x0[0] * x0[0] + x0[1] * x0[1] + x0[2] * x0[2]);
            x0[0] = x0[0] / nrm; x0[1] = x0[1] / nrm; x0[2] = x0[2] / nrm;
        }
        if (n + 1 == next_rec) {
            do {
                record_sample(a, s, p, live, x0[0], x0[1], x0[2], msw, grid_wave());
                ++s;
                next_rec = s < a.S ? a.sched[s] : INT64_MAX;
            } while (next_rec == n + 1);
        }
    }
    if (live) {
        a.mx[p] = x0[0];
        a.my[p] = x0[1];
        a.mz[p] = x0[2];
        if (leader) {
            ia.newton_iters[p] += iters;
            ia.newton_fail[p] += fails;
        }
        uint32_t st = 0;
        if (!(isfinite(x0[0]) && isfinite(x0[1]) && isfinite(x0[2]))) st |= 1u;
        if (fails) st |= 2u;
        a.status[p] |= st;
    }
}

// Sum per-wave partials of each (sample, component) in a fixed order.
static __global__ void ens_reduce_kernel(const double *partial, int n_waves, int64_t rows, double *out)
{
    const int64_t r = blockIdx.x;
    if (r >= rows) return;
    __shared__ double buf[kBlock];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n_waves; i += blockDim.x) acc += partial[r * n_waves + i];
    buf[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) buf[threadIdx.x] += buf[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[r] = buf[0];
}

// normals j0 .. j0+n-1 of one particle's stream, one lane per normal
template <typename T>
__global__ void wiener_kernel(uint32_t key0, uint32_t key1, uint32_t pid, uint32_t tag, uint64_t j0,
                              int64_t n, double *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t j = j0 + (uint64_t)i;
    const uint64_t c = j >> 2;
    PhiloxKeys keys;
    keys.init(key0, key1);
    const u32x4 r = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)(c >> 32), pid, tag}, keys);
    T z[4];
    Num<T>::box_muller(r.x, r.y, z[0], z[1]);
    Num<T>::box_muller(r.z, r.w, z[2], z[3]);
    out[i] = (double)z[j & 3];
}

}  // namespace mp
