// replica_kernels.hpp -- thermal replicas of one cluster geometry.
//
// When every member of an ensemble is the same cluster (same positions,
// radii, anisotropy, axes, material and field) and only the Wiener stream
// (seed) and initial state differ, the dipolar field of all replicas at one
// Heun stage is one matrix product.  With the replica-major state
// Mt[r][3j+b] and the dipole tensor D[3i+a][3j+b] = w_j r_ij^-3 (3 u_a u_b -
// delta_ab) (lib/field.cpp:187-225; zero 3x3 blocks on the diagonal),
//   Ht[r][3i+a] = sum_k Mt[r][k] D[3i+a][k]
// i.e. Ht = Mt D^T.  Both MFMA operands are then "one row, consecutive k"
// reads of row-major matrices, and a wave's output lanes write consecutive
// columns of Ht.
//
// The product runs on MFMA in BF16 with the three-term split
// x = x_hi + x_lo (x_hi = bf16(x), x_lo = bf16(x - x_hi)):
//   Mt D^T ~ Mt_hi D_hi^T + Mt_hi D_lo^T + Mt_lo D_hi^T, FP32 accumulation,
// or in FP32 on MFMA.  The Heun stages around it are elementwise kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sllg_kernels.hpp"

namespace mp {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(16))) float f32x16;

constexpr int kGemmTile = 128;  // block tile (rows of Mt) x (rows of D); 4 waves of 64 x 64

__device__ inline void split_bf16(float x, __bf16 &hi, __bf16 &lo)
{
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// Store one 32x32 MFMA accumulator: element (row, col) of the tile is in lane
// col + 32 * ((row / 4) % 2), register 4 * (row / 8) + row % 4.
__device__ inline void store_tile32(float *C, int ldc, int r0, int c0, const f32x16 &acc)
{
    const int l = threadIdx.x & 63;
    const int col = c0 + (l & 31);
#pragma unroll
    for (int reg = 0; reg < 16; ++reg) {
        const int row = r0 + 8 * (reg >> 2) + 4 * (l >> 5) + (reg & 3);
        C[(size_t)row * ldc + col] = acc[reg];
    }
}

// Ht = Mt_hi D_hi^T + Mt_hi D_lo^T + Mt_lo D_hi^T.  A = Mt [rows][K],
// B = D [cols][K], both row-major BF16; C [rows][ldc] FP32.  rows, cols are
// multiples of kGemmTile and K of 16 (the host pads with zeros).
static __global__ __launch_bounds__(256) void gemm_bf16x3_kernel(const __bf16 *__restrict__ a_hi,
                                                          const __bf16 *__restrict__ a_lo,
                                                          const __bf16 *__restrict__ b_hi,
                                                          const __bf16 *__restrict__ b_lo,
                                                          float *__restrict__ c, int K, int ldc)
{
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int r0 = blockIdx.x * kGemmTile + (wave >> 1) * 64;
    const int c0 = blockIdx.y * kGemmTile + (wave & 1) * 64;
    const int lr = l & 31, kq = 8 * (l >> 5);
    f32x16 acc[2][2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[t][u][q] = 0.f;
    const bf16x8 *ah[2], *al[2], *bh[2], *bl[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const size_t ra = (size_t)(r0 + 32 * t + lr) * K + kq;
        const size_t rb = (size_t)(c0 + 32 * t + lr) * K + kq;
        ah[t] = reinterpret_cast<const bf16x8 *>(a_hi + ra);
        al[t] = reinterpret_cast<const bf16x8 *>(a_lo + ra);
        bh[t] = reinterpret_cast<const bf16x8 *>(b_hi + rb);
        bl[t] = reinterpret_cast<const bf16x8 *>(b_lo + rb);
    }
    for (int k = 0; k < K; k += 16) {
        const int o = k / 8;  // bf16x8 elements
        bf16x8 xa_h[2], xa_l[2], xb_h[2], xb_l[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            xa_h[t] = ah[t][o];
            xa_l[t] = al[t][o];
            xb_h[t] = bh[t][o];
            xb_l[t] = bl[t][o];
        }
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xa_h[t], xb_h[u], acc[t][u], 0, 0, 0);
                acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xa_h[t], xb_l[u], acc[t][u], 0, 0, 0);
                acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xa_l[t], xb_h[u], acc[t][u], 0, 0, 0);
            }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) store_tile32(c, ldc, r0 + 32 * t, c0 + 32 * u, acc[t][u]);
}

// The same product in FP32 on MFMA (v_mfma_f32_32x32x2f32).  Each lane reads
// four consecutive k of its rows per step of 8 and feeds them to four
// instructions; A and B use the same k assignment, so the sum is complete.
static __global__ __launch_bounds__(256) void gemm_f32_kernel(const float *__restrict__ a,
                                                       const float *__restrict__ b,
                                                       float *__restrict__ c, int K, int ldc)
{
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int r0 = blockIdx.x * kGemmTile + (wave >> 1) * 64;
    const int c0 = blockIdx.y * kGemmTile + (wave & 1) * 64;
    const int lr = l & 31, kq = 4 * (l >> 5);
    f32x16 acc[2][2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[t][u][q] = 0.f;
    const f32x4 *pa[2], *pb[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        pa[t] = reinterpret_cast<const f32x4 *>(a + (size_t)(r0 + 32 * t + lr) * K + kq);
        pb[t] = reinterpret_cast<const f32x4 *>(b + (size_t)(c0 + 32 * t + lr) * K + kq);
    }
    for (int k = 0; k < K; k += 8) {
        const int o = k / 4;
        f32x4 xa[2], xb[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            xa[t] = pa[t][o];
            xb[t] = pb[t][o];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 2; ++u)
                    acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x2f32(xa[t][q], xb[u][q], acc[t][u], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) store_tile32(c, ldc, r0 + 32 * t, c0 + 32 * u, acc[t][u]);
}

// FP64 product on the VALU, one output per lane (reference for tests).
static __global__ void gemm_f64_kernel(const double *__restrict__ a, const double *__restrict__ b,
                                double *__restrict__ c, int rows, int cols, int K)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const int r = i / cols, col = i % cols;
    double s = 0.0;
    for (int k = 0; k < K; ++k) s += a[(size_t)r * K + k] * b[(size_t)col * K + k];
    c[(size_t)r * cols + col] = s;
}

// FP32 -> (hi, lo) BF16 split of n values.
static __global__ void split_kernel(const float *__restrict__ x, __bf16 *__restrict__ hi,
                             __bf16 *__restrict__ lo, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    __bf16 h, l;
    split_bf16(x[i], h, l);
    hi[i] = h;
    lo[i] = l;
}

// ------------------------------------------------------------------ Heun stages
// Per-replica Heun stages around the product.  Thread = (replica r, particle
// p), p fastest; state rows are replica-major with leading dimension ld
// (>= 3N, padded).  The applied field of the two stage times is uniform and
// passed in (h1, h2).
struct ReplicaArgs {
    int n;                  // particles per replica
    int reps;               // replicas
    int ld;                 // padded row length (>= 3n)
    int n_waves;            // waves in the grid (ensemble partial stride)
    const float *kex, *key, *kez, *edx, *edy, *edz;  // per particle: k e, e dt
    const float *sig;       // per particle: sr sqrt(dt)
    float dt, alpha;
    const uint32_t *key0, *key1;   // per replica Philox key (member seed)
    const double *msw;             // per replica Ms
    uint32_t tag;
    int renorm;
    float *m, *mt, *delta, *h;     // [reps][ld]: state, predictor, stage-1 increment, field
    __bf16 *m_hi, *m_lo;           // split of the state the next product reads
    uint32_t *status;              // [reps * n]
    // sampling
    const int64_t *sched;
    int64_t S;
    float *samples;                // [S][3][reps * n] or null
    double *ens_partial;           // [S][3][n_waves] or null
};

__device__ inline void replica_normals(const ReplicaArgs &a, int r, int p, int64_t n, float sig,
                                       float &sx, float &sy, float &sz)
{
    PhiloxKeys keys;
    keys.init(a.key0[r], a.key1[r]);
    float z[8];
    const uint64_t c0 = (3 * (uint64_t)n) >> 2;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const uint64_t c = c0 + i;
        const u32x4 w = philox4x32_10(u32x4{(uint32_t)c, (uint32_t)(c >> 32), (uint32_t)p, a.tag}, keys);
        Num<float>::box_muller(w.x, w.y, z[4 * i], z[4 * i + 1]);
        Num<float>::box_muller(w.z, w.w, z[4 * i + 2], z[4 * i + 3]);
    }
    const int o = (int)((3 * (uint64_t)n) & 3);  // normals 3n..3n+2 start at z[o]
    float w0 = z[0], w1 = z[1], w2 = z[2];
    if (o == 1) { w0 = z[1]; w1 = z[2]; w2 = z[3]; }
    else if (o == 2) { w0 = z[2]; w1 = z[3]; w2 = z[4]; }
    else if (o == 3) { w0 = z[3]; w1 = z[4]; w2 = z[5]; }
    sx = sig * w0;
    sy = sig * w1;
    sz = sig * w2;
}

// stage 1: m~ = m + L(m, h dt + s), stores m~, the increment and the split of m~
static __global__ __launch_bounds__(256) void replica_stage1_kernel(ReplicaArgs a, int64_t n_step, float h1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n * a.reps) return;
    const int r = i / a.n, p = i % a.n;
    const size_t o = (size_t)r * a.ld + 3 * p;
    const float mx = a.m[o], my = a.m[o + 1], mz = a.m[o + 2];
    const float hx = a.h[o], hy = a.h[o + 1], hz = a.h[o + 2];
    float sx, sy, sz;
    replica_normals(a, r, p, n_step, a.sig[p], sx, sy, sz);
    const float dt = a.dt;
    const float d = mx * a.kex[p] + my * a.key[p] + mz * a.kez[p];
    const float vx = d * a.edx[p] + (hx * dt + sx), vy = d * a.edy[p] + (hy * dt + sy),
                vz = d * a.edz[p] + ((hz + h1) * dt + sz);
    float ax, ay, az;
    llg_map(mx, my, mz, vx, vy, vz, a.alpha, ax, ay, az);
    const float tx = mx + ax, ty = my + ay, tz = mz + az;
    a.delta[o] = ax; a.delta[o + 1] = ay; a.delta[o + 2] = az;
    a.mt[o] = tx; a.mt[o + 1] = ty; a.mt[o + 2] = tz;
    split_bf16(tx, a.m_hi[o], a.m_lo[o]);
    split_bf16(ty, a.m_hi[o + 1], a.m_lo[o + 1]);
    split_bf16(tz, a.m_hi[o + 2], a.m_lo[o + 2]);
}

// stage 2: m' = m + (L(m, v1) + L(m~, v2)) / 2, stores m', its split, samples
static __global__ __launch_bounds__(256) void replica_stage2_kernel(ReplicaArgs a, int64_t n_step, float h2,
                                                             int64_t s_first, int n_rec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < a.n * a.reps;
    const int q = live ? i : a.n * a.reps - 1;
    const int r = q / a.n, p = q % a.n;
    const size_t o = (size_t)r * a.ld + 3 * p;
    const float tx = a.mt[o], ty = a.mt[o + 1], tz = a.mt[o + 2];
    const float hx = a.h[o], hy = a.h[o + 1], hz = a.h[o + 2];
    float sx, sy, sz;
    replica_normals(a, r, p, n_step, a.sig[p], sx, sy, sz);
    const float dt = a.dt;
    const float d = tx * a.kex[p] + ty * a.key[p] + tz * a.kez[p];
    const float vx = d * a.edx[p] + (hx * dt + sx), vy = d * a.edy[p] + (hy * dt + sy),
                vz = d * a.edz[p] + ((hz + h2) * dt + sz);
    float bx, by, bz;
    llg_map(tx, ty, tz, vx, vy, vz, a.alpha, bx, by, bz);
    float mx = a.m[o] + 0.5f * (a.delta[o] + bx);
    float my = a.m[o + 1] + 0.5f * (a.delta[o + 1] + by);
    float mz = a.m[o + 2] + 0.5f * (a.delta[o + 2] + bz);
    if (a.renorm) {
        const float nrm = sqrtf(mx * mx + my * my + mz * mz);
        mx = mx / nrm;
        my = my / nrm;
        mz = mz / nrm;
    }
    if (live) {
        a.m[o] = mx; a.m[o + 1] = my; a.m[o + 2] = mz;
        split_bf16(mx, a.m_hi[o], a.m_lo[o]);
        split_bf16(my, a.m_hi[o + 1], a.m_lo[o + 1]);
        split_bf16(mz, a.m_hi[o + 2], a.m_lo[o + 2]);
        if (!(isfinite(mx) && isfinite(my) && isfinite(mz))) a.status[i] |= 1u;
    }
    // samples s_first .. s_first + n_rec - 1 are all recorded after this step
    for (int k = 0; k < n_rec; ++k) {
        const int64_t s = s_first + k;
        if (a.samples != nullptr && live) {
            const size_t P = (size_t)a.n * a.reps;
            float *base = a.samples + (size_t)s * 3 * P;
            base[i] = mx;
            base[P + i] = my;
            base[2 * P + i] = mz;
        }
        if (a.ens_partial != nullptr) {
            double x = live ? (double)mx * a.msw[r] : 0.0;
            double y = live ? (double)my * a.msw[r] : 0.0;
            double z = live ? (double)mz * a.msw[r] : 0.0;
            wave_sum3<float>(x, y, z);
            if ((threadIdx.x & (kWave - 1)) == 0) {
                const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
                double *dst = a.ens_partial + (size_t)s * 3 * (size_t)a.n_waves;
                dst[wave] = x;
                dst[(size_t)a.n_waves + wave] = y;
                dst[2 * (size_t)a.n_waves + wave] = z;
            }
        }
    }
}

}  // namespace mp
