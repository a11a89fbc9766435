// adjoint_kernels.hpp -- pathwise sensitivities of the Heun ensemble.
//
// Independent particles (one per member, k_red = v_red = 1), fixed Wiener
// increments.  Step n of a particle (integrator::heun, lib/integrators.cpp:
// 372-405, at t = (n+1) dt as lib/simulation.cpp:355, 372-375 calls it):
//   v1 = dt h(m, s_{n+1}) + sig w_n,     m~ = m + L(m, v1)
//   v2 = dt h(m~, s_{n+2}) + sig w_n,    m' = m + (L(m, v1) + L(m~, v2)) / 2
//   h(m, s) = (m.e) e + happ s z          (lib/field.cpp:37-40, 116-132)
//   L(m, v) = -m x v + alpha (v |m|^2 - m (m.v))   (lib/llg.cpp:14-29, 92-106)
//   s_k = sin(2 pi frac(k f dt)),  f dt the field cycles per step in SI units,
// so the phase does not move with the parameters.
//
// The forward pass stores the state every C steps (checkpoints, [ck][3][P]).
// The reverse pass walks the segments backwards: it recomputes a segment's
// states from its checkpoint into a per-lane scratch ([C][3][P], each lane
// reads back only what it wrote), then carries the adjoint lambda = dJ/dm
// back through the segment and accumulates dJ/d(dt, sig, alpha, happ) in
// FP64.  The objective is linear in the sampled m_z: J = sum_s c_s
// sum_p m_z,p(k_s) + c_f sum_p m_z,p(N), so lambda_z picks up c_s at each
// sample step and c_f at the end.
//
// Transposed derivatives of L used by the reverse step (x a 3-vector):
//   (dL/dv)^T x     = m x x + alpha (|m|^2 x - m (m.x))
//   (dL/dm)^T x     = -v x x + alpha (2 m (v.x) - (m.v) x - v (m.x))
//   dL/dalpha       = v |m|^2 - m (m.v)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sllg_kernels.hpp"

namespace mp {

template <typename T>
struct V3 {
    T x, y, z;
};

template <typename T>
__device__ inline T dot(const V3<T> &a, const V3<T> &b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

template <typename T>
__device__ inline V3<T> cross(const V3<T> &a, const V3<T> &b)
{
    return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <typename T>
struct PathArgs {
    int n;                          // particles
    const T *ex, *ey, *ez;          // easy axes
    const T *sig;                   // sr sqrt(dt) per particle
    const uint32_t *key0, *key1;    // Philox key (member seed) per particle
    uint32_t tag;
    T dt, alpha, happ;
    T *mx, *my, *mz;                // state
    const T *ftab;                  // ftab[i] = s_{tab0 + i}
    int64_t tab0;
    int64_t n_steps;                // N
    int64_t ck_every;               // C (a multiple of 4)
    T *ck;                          // [N / C rounded up][3][n]
    T *scratch;                     // [C][3][n]
    uint32_t *status;               // bit 0: non-finite state
    // forward: steps [n0, n1); reverse: segments [seg_lo, seg_hi] downwards
    int64_t n0, n1;
    int64_t seg_lo, seg_hi;
    // sampling (ascending steps in [0, N])
    const int64_t *sched;
    int64_t S;
    int64_t s_next;                 // forward: first sample with sched >= n0
    double *mz_partial;             // forward: [S + 1][n_waves] wave sums (row S: final)
    int n_waves;
    // reverse
    const double *c_sample;         // [S] weight of sum_p m_z,p(k_s)
    double c_final;
    T *lx, *ly, *lz;                // lambda carried between launches
    double *g;                      // [4][n]: dJ/d(dt, sig, alpha, happ) per particle
    int64_t s_last;                 // reverse: last sample with sched <= this launch's end
};

// field h(m, s) and v = dt h + sig w
template <typename T>
__device__ inline V3<T> path_v(const V3<T> &m, const V3<T> &e, const V3<T> &w, T dt, T sig,
                               T happ, T s)
{
    const T me = dot(m, e);
    return V3<T>{dt * (me * e.x) + sig * w.x, dt * (me * e.y) + sig * w.y,
                 dt * (me * e.z + happ * s) + sig * w.z};
}

template <typename T>
__device__ inline V3<T> path_L(const V3<T> &m, const V3<T> &v, T al)
{
    T ax, ay, az;
    llg_map(m.x, m.y, m.z, v.x, v.y, v.z, al, ax, ay, az);
    return V3<T>{ax, ay, az};
}

template <typename T>
__device__ inline V3<T> path_step(const V3<T> &m, const V3<T> &e, const V3<T> &w, T dt, T sig,
                                  T al, T happ, T s1, T s2)
{
    const V3<T> a1 = path_L(m, path_v(m, e, w, dt, sig, happ, s1), al);
    const V3<T> mt{m.x + a1.x, m.y + a1.y, m.z + a1.z};
    const V3<T> a2 = path_L(mt, path_v(mt, e, w, dt, sig, happ, s2), al);
    return V3<T>{m.x + T(0.5) * (a1.x + a2.x), m.y + T(0.5) * (a1.y + a2.y),
                 m.z + T(0.5) * (a1.z + a2.z)};
}

// Unscaled Wiener normals of step group g (steps 4g..4g+3): w of step 4g+i
// is z[3i..3i+2].
template <typename T>
__device__ inline void path_group_normals(int64_t g, uint32_t tag, const PhiloxKeys &keys, T z[12])
{
    group_normals<T>(g, 0u, tag, keys, T(1), z);
}

template <typename T>
__device__ inline V3<T> load3(const T *base, size_t stride, int p)
{
    return V3<T>{base[p], base[stride + p], base[2 * stride + p]};
}

template <typename T>
__device__ inline void store3(T *base, size_t stride, int p, const V3<T> &v)
{
    base[p] = v.x;
    base[stride + p] = v.y;
    base[2 * stride + p] = v.z;
}

template <typename T>
__device__ inline void path_record(const PathArgs<T> &a, int64_t row, bool live, T mz)
{
    double v = live ? (double)mz : 0.0;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
        a.mz_partial[(size_t)row * a.n_waves + wave] = v;
    }
}

// Forward block [n0, n1): checkpoints at multiples of C, wave sums of m_z at
// the sample steps, the final row after step N.  n0 is a multiple of 4.
template <typename T>
__global__ __launch_bounds__(kWave) void path_forward_kernel(PathArgs<T> a)
{
    const int p0 = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = p0 < a.n;
    const int p = live ? p0 : a.n - 1;
    const size_t P = (size_t)a.n;
    V3<T> m{a.mx[p], a.my[p], a.mz[p]};
    const V3<T> e{a.ex[p], a.ey[p], a.ez[p]};
    const T sig = a.sig[p], dt = a.dt, al = a.alpha, ha = a.happ;
    PhiloxKeys keys;
    keys.init(a.key0[p], a.key1[p]);
    int64_t s = a.s_next;
    // next checkpoint step and its index (n0 need not be a checkpoint)
    int64_t ck_i = (a.n0 + a.ck_every - 1) / a.ck_every;
    int64_t ck_n = ck_i * a.ck_every;
    T z[12];
    for (int64_t n = a.n0; n < a.n1; ++n) {
        if (n == ck_n) {
            if (live) store3(a.ck + (size_t)ck_i * 3 * P, P, p, m);
            ++ck_i;
            ck_n += a.ck_every;
        }
        while (s < a.S && a.sched[s] == n) path_record(a, s++, live, m.z);
        const int i = (int)(n & 3);
        if (i == 0 || n == a.n0) path_group_normals<T>(n >> 2, a.tag, keys, z);
        T w0 = z[0], w1 = z[1], w2 = z[2];
        if (i == 1) { w0 = z[3]; w1 = z[4]; w2 = z[5]; }
        else if (i == 2) { w0 = z[6]; w1 = z[7]; w2 = z[8]; }
        else if (i == 3) { w0 = z[9]; w1 = z[10]; w2 = z[11]; }
        const T s1 = a.ftab[n + 1 - a.tab0], s2 = a.ftab[n + 2 - a.tab0];
        m = path_step(m, e, V3<T>{w0, w1, w2}, dt, sig, al, ha, s1, s2);
    }
    if (a.n1 == a.n_steps) {
        while (s < a.S && a.sched[s] == a.n_steps) path_record(a, s++, live, m.z);
        path_record(a, a.S, live, m.z);
    }
    if (live) {
        a.mx[p] = m.x;
        a.my[p] = m.y;
        a.mz[p] = m.z;
        if (!(isfinite(m.x) && isfinite(m.y) && isfinite(m.z))) a.status[p] |= 1u;
    }
}

// Reverse sweep over segments seg_hi, seg_hi - 1, ..., seg_lo.  On entry
// lambda (lx, ly, lz) is dJ/dm at the end of segment seg_hi, including the
// weights of samples at that step.
template <typename T>
__global__ __launch_bounds__(kWave) void path_reverse_kernel(PathArgs<T> a)
{
    const int p0 = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = p0 < a.n;
    const int p = live ? p0 : a.n - 1;
    const size_t P = (size_t)a.n;
    const V3<T> e{a.ex[p], a.ey[p], a.ez[p]};
    const T sig = a.sig[p], dt = a.dt, al = a.alpha, ha = a.happ;
    PhiloxKeys keys;
    keys.init(a.key0[p], a.key1[p]);
    V3<T> lam{a.lx[p], a.ly[p], a.lz[p]};
    double g_dt = 0.0, g_sig = 0.0, g_al = 0.0, g_ha = 0.0;
    int64_t s = a.s_last;   // samples [0, s] have sched <= current step
    T z[12];
    for (int64_t seg = a.seg_hi; seg >= a.seg_lo; --seg) {
        const int64_t n_a = seg * a.ck_every;
        const int64_t n_b = n_a + a.ck_every < a.n_steps ? n_a + a.ck_every : a.n_steps;
        // 1. recompute the segment's states m_{n_a} .. m_{n_b - 1}
        V3<T> m = load3(a.ck + (size_t)seg * 3 * P, P, p);
        for (int64_t n = n_a; n < n_b; ++n) {
            const int i = (int)(n & 3);
            if (i == 0) path_group_normals<T>(n >> 2, a.tag, keys, z);
            T w0 = z[0], w1 = z[1], w2 = z[2];
            if (i == 1) { w0 = z[3]; w1 = z[4]; w2 = z[5]; }
            else if (i == 2) { w0 = z[6]; w1 = z[7]; w2 = z[8]; }
            else if (i == 3) { w0 = z[9]; w1 = z[10]; w2 = z[11]; }
            if (live) store3(a.scratch + (size_t)(n - n_a) * 3 * P, P, p, m);
            const T s1 = a.ftab[n + 1 - a.tab0], s2 = a.ftab[n + 2 - a.tab0];
            m = path_step(m, e, V3<T>{w0, w1, w2}, dt, sig, al, ha, s1, s2);
        }
        // 2. carry lambda back through steps n_b - 1 .. n_a
        int64_t have = -1;
        for (int64_t n = n_b - 1; n >= n_a; --n) {
            if ((n >> 2) != have) {
                have = n >> 2;
                path_group_normals<T>(have, a.tag, keys, z);
            }
            const int i = (int)(n & 3);
            T w0 = z[0], w1 = z[1], w2 = z[2];
            if (i == 1) { w0 = z[3]; w1 = z[4]; w2 = z[5]; }
            else if (i == 2) { w0 = z[6]; w1 = z[7]; w2 = z[8]; }
            else if (i == 3) { w0 = z[9]; w1 = z[10]; w2 = z[11]; }
            const V3<T> w{w0, w1, w2};
            // padding lanes (p0 >= n) read particle n - 1's slot; their results are dropped
            const V3<T> mm = load3(a.scratch + (size_t)(n - n_a) * 3 * P, P, p);
            const T s1 = a.ftab[n + 1 - a.tab0], s2 = a.ftab[n + 2 - a.tab0];
            // forward intermediates
            const V3<T> v1 = path_v(mm, e, w, dt, sig, ha, s1);
            const V3<T> a1 = path_L(mm, v1, al);
            const V3<T> mt{mm.x + a1.x, mm.y + a1.y, mm.z + a1.z};
            const V3<T> v2 = path_v(mt, e, w, dt, sig, ha, s2);
            // m' = m + (a1 + a2) / 2
            const V3<T> ab{T(0.5) * lam.x, T(0.5) * lam.y, T(0.5) * lam.z};
            // a2 = L(mt, v2)
            const T tt = dot(mt, mt), tv2 = dot(mt, v2);
            const T v2b = dot(v2, ab), tb = dot(mt, ab);
            const V3<T> vx2 = cross(v2, ab), tx2 = cross(mt, ab);
            V3<T> mtb{-vx2.x + al * (T(2) * mt.x * v2b - tv2 * ab.x - v2.x * tb),
                      -vx2.y + al * (T(2) * mt.y * v2b - tv2 * ab.y - v2.y * tb),
                      -vx2.z + al * (T(2) * mt.z * v2b - tv2 * ab.z - v2.z * tb)};
            const V3<T> v2bar{tx2.x + al * (tt * ab.x - mt.x * tb),
                              tx2.y + al * (tt * ab.y - mt.y * tb),
                              tx2.z + al * (tt * ab.z - mt.z * tb)};
            g_al += (double)(dot(v2, ab) * tt - tb * tv2);
            // v2 = dt h(mt, s2) + sig w
            const T ev2 = dot(e, v2bar);
            const T mte = dot(mt, e);
            mtb.x += dt * e.x * ev2;
            mtb.y += dt * e.y * ev2;
            mtb.z += dt * e.z * ev2;
            g_dt += (double)(mte * ev2 + ha * s2 * v2bar.z);
            g_sig += (double)dot(w, v2bar);
            g_ha += (double)(dt * s2 * v2bar.z);
            // mt = m + a1
            const V3<T> a1b{ab.x + mtb.x, ab.y + mtb.y, ab.z + mtb.z};
            V3<T> mb{lam.x + mtb.x, lam.y + mtb.y, lam.z + mtb.z};
            // a1 = L(m, v1)
            const T mm2 = dot(mm, mm), mv1 = dot(mm, v1);
            const T v1b = dot(v1, a1b), mb1 = dot(mm, a1b);
            const V3<T> vx1 = cross(v1, a1b), mx1 = cross(mm, a1b);
            mb.x += -vx1.x + al * (T(2) * mm.x * v1b - mv1 * a1b.x - v1.x * mb1);
            mb.y += -vx1.y + al * (T(2) * mm.y * v1b - mv1 * a1b.y - v1.y * mb1);
            mb.z += -vx1.z + al * (T(2) * mm.z * v1b - mv1 * a1b.z - v1.z * mb1);
            const V3<T> v1bar{mx1.x + al * (mm2 * a1b.x - mm.x * mb1),
                              mx1.y + al * (mm2 * a1b.y - mm.y * mb1),
                              mx1.z + al * (mm2 * a1b.z - mm.z * mb1)};
            g_al += (double)(v1b * mm2 - mb1 * mv1);
            // v1 = dt h(m, s1) + sig w
            const T ev1 = dot(e, v1bar);
            const T me = dot(mm, e);
            mb.x += dt * e.x * ev1;
            mb.y += dt * e.y * ev1;
            mb.z += dt * e.z * ev1;
            g_dt += (double)(me * ev1 + ha * s1 * v1bar.z);
            g_sig += (double)dot(w, v1bar);
            g_ha += (double)(dt * s1 * v1bar.z);
            lam = mb;
            // weights of samples taken at step n
            while (s >= 0 && a.sched[s] == n) lam.z += (T)a.c_sample[s--];
        }
    }
    if (live) {
        a.lx[p] = lam.x;
        a.ly[p] = lam.y;
        a.lz[p] = lam.z;
        a.g[p] += g_dt;
        a.g[P + p] += g_sig;
        a.g[2 * P + p] += g_al;
        a.g[3 * P + p] += g_ha;
        if (!(isfinite(lam.x) && isfinite(lam.y) && isfinite(lam.z))) a.status[p] |= 2u;
    }
}

// Field table: out[i] = sin(2 pi frac((base + i) fdt)), FP64 phase reduction.
template <typename T>
__global__ void path_field_table_kernel(double fdt, int64_t base, int64_t count, T *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double c = (double)(base + i) * fdt;
    c -= floor(c);
    out[i] = (T)sin(2.0 * kPi * c);
}

}  // namespace mp
