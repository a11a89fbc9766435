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
// The forward pass stores the state every C steps (checkpoints).  The
// objective is linear in the sampled m_z: J = sum_s c_s sum_p m_z,p(k_s) +
// c_f sum_p m_z,p(N), so the adjoint lambda = dJ/dm picks up c_s at each
// sample step and c_f at the end, and between sample steps it is a linear
// map of its end value.  Those maps (transfers, below) are computed right
// after each forward batch, for all segments of the batch in parallel: a
// lane per (segment, particle) recomputes its segment from the checkpoint
// into scratch ([C][6][lanes]: states and normals, read back only by the lane
// that wrote it) and carries three basis adjoints back through it.  They are folded into one
// transfer per sampling interval, and a reverse pass with given weights is
// then a short scan over the intervals.
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
    T *ck;                          // checkpoints [segment - ck_base][3][n]
    int64_t ck_base;
    uint32_t *status;               // bit 0: non-finite state
    int64_t n0, n1;                 // steps [n0, n1) of this launch
    // sampling (ascending steps in [0, N])
    const int64_t *sched;
    int64_t S;
    int64_t s_next;                 // first sample with sched >= n0
    double *mz_partial;             // [S + 1][n_waves] wave sums (row S: final)
    int n_waves;
};

// w of step n from the unscaled normals of its 4-step group (no dynamic
// register indexing)
template <typename T>
__device__ inline V3<T> group_w(const T z[12], int i)
{
    if (i == 0) return V3<T>{z[0], z[1], z[2]};
    if (i == 1) return V3<T>{z[3], z[4], z[5]};
    if (i == 2) return V3<T>{z[6], z[7], z[8]};
    return V3<T>{z[9], z[10], z[11]};
}

// field h(m, s) and v = dt h + sig w (fused multiply-adds fixed by the
// source, as in llg_map, so the recompute reproduces the forward pass)
template <typename T>
__device__ inline V3<T> path_v(const V3<T> &m, const V3<T> &e, const V3<T> &w, T dt, T sig,
                               T happ, T s)
{
#pragma clang fp contract(off)
    const T me = fmad(m.x, e.x, fmad(m.y, e.y, m.z * e.z));
    return V3<T>{fmad(dt, me * e.x, sig * w.x), fmad(dt, me * e.y, sig * w.y),
                 fmad(dt, fmad(me, e.z, happ * s), sig * w.z)};
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
#pragma clang fp contract(off)
    const V3<T> a1 = path_L(m, path_v(m, e, w, dt, sig, happ, s1), al);
    const V3<T> mt{m.x + a1.x, m.y + a1.y, m.z + a1.z};
    const V3<T> a2 = path_L(mt, path_v(mt, e, w, dt, sig, happ, s2), al);
    return V3<T>{fmad(T(0.5), a1.x + a2.x, m.x), fmad(T(0.5), a1.y + a2.y, m.y),
                 fmad(T(0.5), a1.z + a2.z, m.z)};
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

// wave sum of m_z into row `row` of the partial sums; `wave` indexes the
// 64 particles of the calling wave
template <typename T>
__device__ inline void path_record(const PathArgs<T> &a, int64_t row, int wave, bool live, T mz)
{
    double v = live ? (double)mz : 0.0;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) a.mz_partial[(size_t)row * a.n_waves + wave] = v;
}

// Step groups per half of the forward kernel's normal ring
template <typename T>
struct FwdRing {
    static constexpr int kGroups = sizeof(T) == 4 ? 8 : 4;
};

// Forward block [n0, n1): checkpoints at multiples of C, wave sums of m_z at
// the sample steps, the final row after step N.  n0 is a multiple of 4 (and
// of C).  A workgroup of two waves serves 64 particles: wave 1 generates the
// Wiener normals of 4-step groups into an LDS ring of two halves, wave 0
// integrates from the other half, and the two swap halves at a barrier every
// kGroups groups.  The integrating wave thus carries only the Heun steps; it
// reads the group's five field values once and keeps the next sample step in
// a register.
template <typename T>
__global__ __launch_bounds__(2 * kWave) void path_forward_kernel(PathArgs<T> a)
{
    constexpr int K = FwdRing<T>::kGroups;
    __shared__ T ring[2][K][12][kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const bool producer = __builtin_amdgcn_readfirstlane(threadIdx.x) >= kWave;
    const int wave = blockIdx.x;
    const int p0 = wave * kWave + lane;
    const bool live = p0 < a.n;
    const int p = live ? p0 : a.n - 1;
    const size_t P = (size_t)a.n;
    const int64_t g0 = a.n0 >> 2, g1 = (a.n1 + 3) >> 2;
    const int64_t n_chunks = (g1 - g0 + K - 1) / K;
    PhiloxKeys keys;
    keys.init(a.key0[p], a.key1[p]);
    V3<T> m{a.mx[p], a.my[p], a.mz[p]};
    const V3<T> e{a.ex[p], a.ey[p], a.ez[p]};
    const T sig = a.sig[p], dt = a.dt, al = a.alpha, ha = a.happ;
    int64_t s = a.s_next;
    int64_t next_rec = s < a.S ? a.sched[s] : INT64_MAX;
    int64_t ck_n = (a.n0 + a.ck_every - 1) / a.ck_every * a.ck_every;
    if (!producer) {
        // the integrating wave takes issue priority over co-resident waves
        __builtin_amdgcn_s_setprio(3);
    }
    // pass c: the producer fills chunk c, the integrator runs chunk c - 1
    for (int64_t c = 0; c <= n_chunks; ++c) {
        if (producer) {
            if (c < n_chunks) {
                T (*half)[12][kWave] = ring[c & 1];
                for (int k = 0; k < K; ++k) {
                    const int64_t g = g0 + c * K + k;
                    if (g >= g1) break;
                    T z[12];
                    path_group_normals<T>(g, a.tag, keys, z);
#pragma unroll
                    for (int q = 0; q < 12; ++q) half[k][q][lane] = z[q];
                }
            }
        } else if (c > 0) {
            T (*half)[12][kWave] = ring[(c - 1) & 1];
            for (int k = 0; k < K; ++k) {
                const int64_t g = g0 + (c - 1) * K + k;
                if (g >= g1) break;
                T zc[12];
#pragma unroll
                for (int q = 0; q < 12; ++q) zc[q] = half[k][q][lane];
                const int64_t nb = 4 * g;
                const T *tab = a.ftab + (nb + 1 - a.tab0);
                const T f0 = tab[0], f1 = tab[1], f2 = tab[2], f3 = tab[3], f4 = tab[4];
                if (nb == ck_n) {
                    if (live) store3(a.ck + (size_t)(nb / a.ck_every - a.ck_base) * 3 * P, P, p, m);
                    ck_n += a.ck_every;
                }
                if (nb + 4 <= a.n1 && next_rec >= nb + 4) {
                    m = path_step(m, e, V3<T>{zc[0], zc[1], zc[2]}, dt, sig, al, ha, f0, f1);
                    m = path_step(m, e, V3<T>{zc[3], zc[4], zc[5]}, dt, sig, al, ha, f1, f2);
                    m = path_step(m, e, V3<T>{zc[6], zc[7], zc[8]}, dt, sig, al, ha, f2, f3);
                    m = path_step(m, e, V3<T>{zc[9], zc[10], zc[11]}, dt, sig, al, ha, f3, f4);
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int64_t n = nb + i;
                        if (n >= a.n1) break;
                        while (n == next_rec) {
                            path_record(a, s++, wave, live, m.z);
                            next_rec = s < a.S ? a.sched[s] : INT64_MAX;
                        }
                        const T sa = i == 0 ? f0 : i == 1 ? f1 : i == 2 ? f2 : f3;
                        const T sb = i == 0 ? f1 : i == 1 ? f2 : i == 2 ? f3 : f4;
                        m = path_step(m, e, group_w(zc, i), dt, sig, al, ha, sa, sb);
                    }
                }
            }
        }
        __syncthreads();
    }
    if (producer) return;
    if (a.n1 == a.n_steps) {
        while (next_rec == a.n_steps) {
            path_record(a, s++, wave, live, m.z);
            next_rec = s < a.S ? a.sched[s] : INT64_MAX;
        }
        path_record(a, a.S, wave, live, m.z);
    }
    if (live) {
        a.mx[p] = m.x;
        a.my[p] = m.y;
        a.mz[p] = m.z;
        if (!(isfinite(m.x) && isfinite(m.y) && isfinite(m.z))) a.status[p] |= 1u;
    }
}

// ------------------------------------------------------------------ reverse
// The adjoint is linear in lambda: over any stretch [x, y) of steps
//   lambda_x = Phi^T lambda_y,   g += G^T lambda_y
// with Phi^T (3 x 3) and G^T (4 x 3) independent of the objective's weights.
// The transfer kernel computes them for every stretch between consecutive
// sample steps inside each segment, all segments of a batch in parallel (one
// lane per (segment, particle)); the compose kernel folds the stretches of one
// sampling interval into one transfer; the scan kernel runs the weighted
// reverse pass over the intervals.  Transfers compose as
//   (Phi_A^T, G_A^T) then (Phi_B^T, G_B^T)  ->  (Phi_A^T Phi_B^T, G_A^T Phi_B^T + G_B^T)
// for A = [x, y) followed by B = [y, z).  A transfer is 21 doubles:
// Phi^T row-major (9), then G^T row-major (12).
constexpr int kTransfer = 21;

// Forward intermediates of one step needed by its adjoint
template <typename T>
struct StepFwd {
    V3<T> m, v1, mt, v2, w;
    T s1, s2;
};

template <typename T>
__device__ inline StepFwd<T> step_forward(const V3<T> &m, const V3<T> &e, const V3<T> &w, T dt,
                                          T sig, T al, T ha, T s1, T s2)
{
#pragma clang fp contract(off)
    StepFwd<T> f;
    f.m = m;
    f.w = w;
    f.s1 = s1;
    f.s2 = s2;
    f.v1 = path_v(m, e, w, dt, sig, ha, s1);
    const V3<T> a1 = path_L(m, f.v1, al);
    f.mt = V3<T>{m.x + a1.x, m.y + a1.y, m.z + a1.z};
    f.v2 = path_v(f.mt, e, w, dt, sig, ha, s2);
    return f;
}

// lambda_n = (dm_{n+1}/dm_n)^T lambda_{n+1}, and g += (dm_{n+1}/d(dt, sig,
// alpha, happ))^T lambda_{n+1}, for m_{n+1} = m + (L(m, v1) + L(mt, v2)) / 2.
template <typename T>
__device__ inline V3<T> step_adjoint(const StepFwd<T> &f, const V3<T> &e, T dt, T al, T ha,
                                     const V3<T> &lam, T g[4])
{
    const V3<T> &mm = f.m, &v1 = f.v1, &mt = f.mt, &v2 = f.v2, &w = f.w;
    const V3<T> ab{T(0.5) * lam.x, T(0.5) * lam.y, T(0.5) * lam.z};
    // a2 = L(mt, v2)
    const T tt = dot(mt, mt), tv2 = dot(mt, v2);
    const T v2b = dot(v2, ab), tb = dot(mt, ab);
    const V3<T> vx2 = cross(v2, ab), tx2 = cross(mt, ab);
    V3<T> mtb{-vx2.x + al * (T(2) * mt.x * v2b - tv2 * ab.x - v2.x * tb),
              -vx2.y + al * (T(2) * mt.y * v2b - tv2 * ab.y - v2.y * tb),
              -vx2.z + al * (T(2) * mt.z * v2b - tv2 * ab.z - v2.z * tb)};
    const V3<T> v2bar{tx2.x + al * (tt * ab.x - mt.x * tb), tx2.y + al * (tt * ab.y - mt.y * tb),
                      tx2.z + al * (tt * ab.z - mt.z * tb)};
    g[2] += v2b * tt - tb * tv2;
    // v2 = dt h(mt, s2) + sig w
    const T ev2 = dot(e, v2bar);
    const T mte = dot(mt, e);
    mtb.x += dt * e.x * ev2;
    mtb.y += dt * e.y * ev2;
    mtb.z += dt * e.z * ev2;
    g[0] += mte * ev2 + ha * f.s2 * v2bar.z;
    g[1] += dot(w, v2bar);
    g[3] += dt * f.s2 * v2bar.z;
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
    const V3<T> v1bar{mx1.x + al * (mm2 * a1b.x - mm.x * mb1), mx1.y + al * (mm2 * a1b.y - mm.y * mb1),
                      mx1.z + al * (mm2 * a1b.z - mm.z * mb1)};
    g[2] += v1b * mm2 - mb1 * mv1;
    // v1 = dt h(m, s1) + sig w
    const T ev1 = dot(e, v1bar);
    const T me = dot(mm, e);
    mb.x += dt * e.x * ev1;
    mb.y += dt * e.y * ev1;
    mb.z += dt * e.z * ev1;
    g[0] += me * ev1 + ha * f.s1 * v1bar.z;
    g[1] += dot(w, v1bar);
    g[3] += dt * f.s1 * v1bar.z;
    return mb;
}

template <typename T>
struct TransferArgs {
    int n;                          // particles
    const T *ex, *ey, *ez, *sig;
    const uint32_t *key0, *key1;
    uint32_t tag;
    T dt, alpha, happ;
    const T *ftab;                  // ftab[i] = s_{tab0 + i}
    int64_t tab0;
    int64_t n_steps, ck_every;
    const T *ck;                    // checkpoints of this batch's segments [seg][3][n]
    int64_t seg0, n_segs;           // global index of the first segment, segments in batch
    T *scratch;                     // [C][kScratchRows][n_segs * n]
    const int32_t *seg_first_piece, *seg_n_pieces;   // per segment of the batch
    const int64_t *piece_start;     // first step of each piece (run-wide index)
    int64_t piece_base;             // run-wide index of the batch's first piece
    double *pieces;                 // [piece - piece_base][21][n]
};

// Rows of scratch per step: the state m_n (3), then the unscaled normals w_n
// (3), so that the reverse sweep needs no second pass of the generator.
constexpr int kScratchRows = 6;

// One lane per (segment, particle), particle fastest.
// Register budget of four waves per SIMD (128 VGPRs, a few spills outside the
// step body): 4% faster than the three waves the unconstrained kernel gets.
template <typename T>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(4, 8))) void
path_transfer_kernel(TransferArgs<T> a)
{
    const int64_t lanes = a.n_segs * (int64_t)a.n;
    const int64_t gid0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = gid0 < lanes;
    const int64_t gid = live ? gid0 : lanes - 1;
    const int sl = (int)(gid / a.n), p = (int)(gid % a.n);
    const size_t P = (size_t)a.n, L = (size_t)lanes;
    const int64_t seg = a.seg0 + sl;
    const int64_t n_a = seg * a.ck_every;
    const int64_t n_b = n_a + a.ck_every < a.n_steps ? n_a + a.ck_every : a.n_steps;
    const V3<T> e{a.ex[p], a.ey[p], a.ez[p]};
    const T sig = a.sig[p], dt = a.dt, al = a.alpha, ha = a.happ;
    PhiloxKeys keys;
    keys.init(a.key0[p], a.key1[p]);
    T z[12];
    // 1. the segment's states m_{n_a} .. m_{n_b - 1} and their normals
    //    (n_a is a multiple of 4)
    V3<T> m = load3(a.ck + (size_t)sl * 3 * P, P, p);
    for (int64_t n = n_a; n < n_b; ++n) {
        const int i = (int)(n & 3);
        if (i == 0) path_group_normals<T>(n >> 2, a.tag, keys, z);
        const V3<T> w = group_w(z, i);
        if (live) {
            T *row = a.scratch + (size_t)(n - n_a) * kScratchRows * L;
            store3(row, L, (int)gid, m);
            store3(row + 3 * L, L, (int)gid, w);
        }
        m = path_step(m, e, w, dt, sig, al, ha, a.ftab[n + 1 - a.tab0], a.ftab[n + 2 - a.tab0]);
    }
    // 2. three adjoint channels (lambda = e_x, e_y, e_z at the end of each
    //    piece) back through the segment; a piece ends at every sample step.
    //    Parameter terms gather in gs (the state type) and move to g (double)
    //    every 4 steps and at piece ends.
    V3<T> lam[3];
    double g[3][4];
    T gs[3][4];
    auto clear_gs = [&]() {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) gs[c][q] = T(0);
    };
    auto flush = [&]() {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) g[c][q] += (double)gs[c][q];
        clear_gs();
    };
    auto reset = [&]() {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lam[c] = V3<T>{T(c == 0), T(c == 1), T(c == 2)};
#pragma unroll
            for (int q = 0; q < 4; ++q) g[c][q] = 0.0;
        }
        clear_gs();
    };
    reset();
    int piece = a.seg_first_piece[sl] + a.seg_n_pieces[sl] - 1;
    int64_t piece_begin = a.piece_start[piece];
    // the rows of step n - 1 load while step n runs
    const T *row = a.scratch + (size_t)(n_b - 1 - n_a) * kScratchRows * L;
    V3<T> mm_next = load3(row, L, (int)gid), w_next = load3(row + 3 * L, L, (int)gid);
    for (int64_t n = n_b - 1; n >= n_a; --n) {
        const V3<T> mm = mm_next, w = w_next;
        if (n > n_a) {
            row -= kScratchRows * L;
            mm_next = load3(row, L, (int)gid);
            w_next = load3(row + 3 * L, L, (int)gid);
        }
        const StepFwd<T> f = step_forward(mm, e, w, dt, sig, al, ha, a.ftab[n + 1 - a.tab0],
                                          a.ftab[n + 2 - a.tab0]);
#pragma unroll
        for (int c = 0; c < 3; ++c) lam[c] = step_adjoint(f, e, dt, al, ha, lam[c], gs[c]);
        if ((n & 3) == 0) flush();
        if (n == piece_begin) {
            flush();
            if (live) {
                double *out = a.pieces + (size_t)(piece - a.piece_base) * kTransfer * P + p;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    out[(3 * r + 0) * P] = (double)(r == 0 ? lam[0].x : r == 1 ? lam[0].y : lam[0].z);
                    out[(3 * r + 1) * P] = (double)(r == 0 ? lam[1].x : r == 1 ? lam[1].y : lam[1].z);
                    out[(3 * r + 2) * P] = (double)(r == 0 ? lam[2].x : r == 1 ? lam[2].y : lam[2].z);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int c = 0; c < 3; ++c) out[(9 + 3 * q + c) * P] = g[c][q];
            }
            reset();
            if (piece > a.seg_first_piece[sl]) piece_begin = a.piece_start[--piece];
        }
    }
}

// Fold a batch's pieces (forward order) into the running transfer of the
// open sampling interval; a piece that starts an interval first closes the
// previous one into intervals[iv - 1].  acc [21][n] carries the open
// interval between batches.  close_last writes the open interval to
// intervals[last_iv] at the end (the last batch of the run).
struct ComposeArgs {
    int n;
    int n_pieces;
    const double *pieces;           // [piece][21][n]
    const int32_t *piece_new;       // 1: the piece starts interval piece_iv
    const int32_t *piece_iv;
    double *acc;                    // [21][n]
    double *intervals;              // [interval][21][n]
    int close_last;
    int last_iv;
};

__global__ __launch_bounds__(kWave) void path_compose_kernel(ComposeArgs a)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const size_t P = (size_t)a.n;
    double acc[kTransfer];
#pragma unroll
    for (int q = 0; q < kTransfer; ++q) acc[q] = a.acc[q * P + p];
    for (int k = 0; k < a.n_pieces; ++k) {
        double x[kTransfer];
#pragma unroll
        for (int q = 0; q < kTransfer; ++q) x[q] = a.pieces[((size_t)k * kTransfer + q) * P + p];
        if (a.piece_new[k]) {
            if (a.piece_iv[k] > 0)
#pragma unroll
                for (int q = 0; q < kTransfer; ++q)
                    a.intervals[((size_t)(a.piece_iv[k] - 1) * kTransfer + q) * P + p] = acc[q];
#pragma unroll
            for (int q = 0; q < kTransfer; ++q) acc[q] = x[q];
        } else {
            // acc then x:  Phi = Phi_acc Phi_x,  G = G_acc Phi_x + G_x  (transposes)
            double r[kTransfer];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    r[3 * i + j] = acc[3 * i] * x[j] + acc[3 * i + 1] * x[3 + j] + acc[3 * i + 2] * x[6 + j];
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    r[9 + 3 * q + j] = acc[9 + 3 * q] * x[j] + acc[9 + 3 * q + 1] * x[3 + j] +
                                       acc[9 + 3 * q + 2] * x[6 + j] + x[9 + 3 * q + j];
#pragma unroll
            for (int q = 0; q < kTransfer; ++q) acc[q] = r[q];
        }
    }
    if (a.close_last) {
#pragma unroll
        for (int q = 0; q < kTransfer; ++q)
            a.intervals[((size_t)a.last_iv * kTransfer + q) * P + p] = acc[q];
    }
#pragma unroll
    for (int q = 0; q < kTransfer; ++q) a.acc[q * P + p] = acc[q];
}

// Weighted reverse pass over the sampling intervals, last to first:
// g += G_i^T lambda; lambda = Phi_i^T lambda + inj[i] e_z (inj[i]: weights of
// the samples at the interval's first step).
struct ScanArgs {
    int n;
    int n_intervals;
    const double *intervals;        // [interval][21][n]
    const double *inj;              // [interval]
    double lambda_z0;               // weights at step N
    double *g;                      // [4][n] out
};

__global__ __launch_bounds__(kWave) void path_scan_kernel(ScanArgs a)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const size_t P = (size_t)a.n;
    double l0 = 0.0, l1 = 0.0, l2 = a.lambda_z0;
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = a.n_intervals - 1; i >= 0; --i) {
        const double *t = a.intervals + (size_t)i * kTransfer * P + p;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            g[q] += t[(9 + 3 * q) * P] * l0 + t[(9 + 3 * q + 1) * P] * l1 + t[(9 + 3 * q + 2) * P] * l2;
        const double n0 = t[0] * l0 + t[P] * l1 + t[2 * P] * l2;
        const double n1 = t[3 * P] * l0 + t[4 * P] * l1 + t[5 * P] * l2;
        const double n2 = t[6 * P] * l0 + t[7 * P] * l1 + t[8 * P] * l2 + a.inj[i];
        l0 = n0;
        l1 = n1;
        l2 = n2;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) a.g[q * P + p] = g[q];
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
