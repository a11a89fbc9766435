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
// Columns are laid out component-blocked per 32 particles: particle i,
// component a sits in column 96 (i / 32) + 32 a + i % 32 (replica_col), with
// the particle count padded to a multiple of 128 so ld = 3 Np is a multiple
// of 384.  The same permutation is applied to both indices of D, so the
// product is unchanged; a 32-column MFMA tile then holds one component of 32
// consecutive particles, and stage kernels read x, y, z as three coalesced
// 128-byte rows.
//
// The product runs on MFMA in BF16 with the three-term split
// x = x_hi + x_lo (x_hi = bf16(x), x_lo = bf16(x - x_hi)):
//   Mt D^T ~ Mt_hi D_hi^T + Mt_hi D_lo^T + Mt_lo D_hi^T, FP32 accumulation,
// or in FP32 on MFMA.  The BF16x3 product runs the Heun stages in its
// epilogue; after the FP32 product they are elementwise kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "sllg_kernels.hpp"

namespace mp {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(16))) float f32x16;

constexpr int kGemmTile = 128;  // block tile of the FP32 product: 4 waves of 64 x 64
constexpr int kParticlePad = 128;  // particles per replica padded to a multiple of this

__host__ __device__ inline int replica_col(int i, int a) { return 96 * (i >> 5) + 32 * a + (i & 31); }

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

// ------------------------------------------------------------------ Heun stages
// Per-replica Heun stages around the product.  Thread = (replica r, particle
// p), p fastest; state rows are replica-major with leading dimension ld
// (>= 3N, padded).  The applied field of the two stage times is uniform and
// passed in (h1, h2).
struct ReplicaArgs {
    int n;                  // particles per replica
    int reps;               // replicas
    int ld;                 // row length 3 Np (Np = n padded to a multiple of 128)
    int n_waves;            // waves in the grid (ensemble partial stride)
    const float *kex, *key, *kez, *edx, *edy, *edz;  // per particle: k e, e dt
    const float *sig;       // per particle: sr sqrt(dt)
    float dt, alpha;
    const uint32_t *key0, *key1;   // per replica Philox key (member seed)
    const double *msw;             // per replica Ms
    uint32_t tag;
    int renorm;
    float *m, *mt, *h;             // [reps][ld]: state, predictor, field
    float *nz;                     // [reps][ld]: the step's scaled Wiener increments (fused
                                   // product: stage 1 writes them, stage 2 reads them)
    uint32_t *status;              // [reps * n]
    // sampling
    const int64_t *sched;
    int64_t S;
    float *samples;                // [S][3][reps * n] or null
    double *ens_partial;           // [S][3][n_waves] or null
};

// Normals 3n, 3n+1, 3n+2 of particle p's stream under the Philox key
// (k0, k1) of its replica, scaled by sig: normal j is component j % 4 of
// Box-Muller over Philox call j / 4 (the stream of philox.h).  Only the
// Box-Muller pairs that hold the three are formed: the two of call 3n/4 when
// they start at component 0 or 1, else the second pair of that call and the
// first of the next.
__device__ inline void replica_normals(uint32_t k0, uint32_t k1, uint32_t tag, int p, int64_t n,
                                       float sig, float &sx, float &sy, float &sz)
{
    PhiloxKeys keys;
    keys.init(k0, k1);
    const uint64_t j0 = 3 * (uint64_t)n, c0 = j0 >> 2;
    const int o = (int)(j0 & 3);
    const u32x4 w = philox4x32_10(u32x4{(uint32_t)c0, (uint32_t)(c0 >> 32), (uint32_t)p, tag}, keys);
    float z0, z1, z2;
    if (o < 2) {
        float b0, b1, b2, b3;
        Num<float>::box_muller(w.x, w.y, b0, b1);
        Num<float>::box_muller(w.z, w.w, b2, b3);
        z0 = o == 0 ? b0 : b1;
        z1 = o == 0 ? b1 : b2;
        z2 = o == 0 ? b2 : b3;
    } else {
        const uint64_t c1 = c0 + 1;
        const u32x4 w1 =
            philox4x32_10(u32x4{(uint32_t)c1, (uint32_t)(c1 >> 32), (uint32_t)p, tag}, keys);
        float b2, b3, b4, b5;
        Num<float>::box_muller(w.z, w.w, b2, b3);
        Num<float>::box_muller(w1.x, w1.y, b4, b5);
        z0 = o == 2 ? b2 : b3;
        z1 = o == 2 ? b3 : b4;
        z2 = o == 2 ? b4 : b5;
    }
    sx = sig * z0;
    sy = sig * z1;
    sz = sig * z2;
}

// Per-particle constants of the stages: k e, e dt and sr sqrt(dt).
struct ReplicaParticle {
    float kex, key, kez, edx, edy, edz, sig;
};

__device__ inline ReplicaParticle replica_particle(const ReplicaArgs &a, int p)
{
    return ReplicaParticle{a.kex[p], a.key[p], a.kez[p], a.edx[p], a.edy[p], a.edz[p], a.sig[p]};
}

// The stage arithmetic of one (replica, particle) element from values
// already in registers: the field (hx, hy, hz) is the product's output,
// (sx, sy, sz) the step's scaled Wiener increments (replica_normals).  Both
// the fused epilogue and the elementwise kernels evaluate these same
// expressions.
// Stage 1: m~ = m + L(m, h dt + s).
__device__ inline void replica_stage1_math(const ReplicaArgs &a, const ReplicaParticle &c,
                                           float sx, float sy, float sz, float h1, float hx,
                                           float hy, float hz, float mx, float my, float mz,
                                           float &tx, float &ty, float &tz)
{
    const float dt = a.dt;
    const float d = mx * c.kex + my * c.key + mz * c.kez;
    const float vx = d * c.edx + (hx * dt + sx), vy = d * c.edy + (hy * dt + sy),
                vz = d * c.edz + ((hz + h1) * dt + sz);
    float ax, ay, az;
    llg_map(mx, my, mz, vx, vy, vz, a.alpha, ax, ay, az);
    tx = mx + ax;
    ty = my + ay;
    tz = mz + az;
}

// Stage 2: m' = m + (L(m, v1) + L(m~, v2)) / 2 with L(m, v1) = m~ - m,
// renormalised when asked.
__device__ inline void replica_stage2_math(const ReplicaArgs &a, const ReplicaParticle &c,
                                           float sx, float sy, float sz, float h2, float hx,
                                           float hy, float hz, float tx, float ty, float tz,
                                           float m0x, float m0y, float m0z, float &mx, float &my,
                                           float &mz)
{
    const float dt = a.dt;
    const float d = tx * c.kex + ty * c.key + tz * c.kez;
    const float vx = d * c.edx + (hx * dt + sx), vy = d * c.edy + (hy * dt + sy),
                vz = d * c.edz + ((hz + h2) * dt + sz);
    float bx, by, bz;
    llg_map(tx, ty, tz, vx, vy, vz, a.alpha, bx, by, bz);
    mx = m0x + 0.5f * ((tx - m0x) + bx);
    my = m0y + 0.5f * ((ty - m0y) + by);
    mz = m0z + 0.5f * ((tz - m0z) + bz);
    if (a.renorm) {
        const float nrm = sqrtf(mx * mx + my * my + mz * mz);
        mx = mx / nrm;
        my = my / nrm;
        mz = mz / nrm;
    }
}

// Store m' of a live element and flag a non-finite state.
__device__ inline void replica_store_state(const ReplicaArgs &a, int r, int p, size_t ox, float mx,
                                           float my, float mz)
{
    a.m[ox] = mx;
    a.m[ox + 32] = my;
    a.m[ox + 64] = mz;
    if (!(isfinite(mx) && isfinite(my) && isfinite(mz))) a.status[(size_t)r * a.n + p] |= 1u;
}

// Ht = Mt_hi D_hi^T + Mt_hi D_lo^T + Mt_lo D_hi^T with LDS staging; Mt is
// read in FP32 and split on its way to LDS, D comes pre-split.  Block
// tile 128 rows x 192 columns; each tile of 64 x 96 (2 x 3 MFMA tiles) is
// computed by two waves (8 waves per block), one per 16-k half of every
// 32-k stage, so each SIMD holds two waves and one can issue MFMAs while the
// other waits on LDS or the barrier; the halves are added through LDS at the
// end.  Stages are double-buffered in LDS and prefetched three stages ahead
// in registers.  Blocks are numbered so that the blocks one XCD runs (block
// ids b with equal b % 8) share panels: with row_panels, row panels of Mt
// (each XCD reads an eighth of Mt and all of D), else column panels of D;
// the shared panels then stay in that XCD's L2.  rows % 128 == 0,
// cols % 192 == 0, K % 192 == 0 (stages come in sixes).  MFMAs are ordered so that six independent accumulators separate
// dependent ones.
//
// kEpi selects the epilogue: 0 stores the product Ht to c; 1 and 2 run the
// Heun stage 1 or 2 on the tile instead (ra, stage field hstage, step
// n_step; stage 2 records the samples s_first .. s_first + n_rec - 1), so
// the field never leaves the kernel.  A lane holds x, y and z of the same
// (replica, particle) in its three column subtiles (u = 0, 1, 2), so the
// stage is per lane; the two k halves each finish half of the registers
// (kh = 0 the registers 0-7 of every subtile, kh = 1 the registers 8-15).
constexpr int kBM = 128, kBN = 192, kBK = 32;

template <int kEpi>
static __global__ __launch_bounds__(512) void gemm_bf16x3_ksplit_kernel(
    const float *__restrict__ a, const __bf16 *__restrict__ b_hi,
    const __bf16 *__restrict__ b_lo, float *__restrict__ c, int K, int ldc, int n_row_blocks,
    int n_col_blocks, ReplicaArgs ra, int64_t n_step, float hstage, int64_t s_first, int n_rec,
    int row_panels)
{
    constexpr int kLdsRow = kBK + 8;
    constexpr int kPanelA = kBM * kLdsRow, kPanelB = kBN * kLdsRow;   // bf16 per panel
    // one LDS array: [buffer][hi/lo] A panels, then B panels; reused at the
    // end for the reduction of the two k halves
    __shared__ __attribute__((aligned(16))) __bf16 lds[4 * kPanelA + 4 * kPanelB];
    auto sa = [&](int buf, int h) { return lds + (2 * buf + h) * kPanelA; };
    auto sb = [&](int buf, int h) { return lds + 4 * kPanelA + (2 * buf + h) * kPanelB; };
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    const int kh = wave >> 2, w4 = wave & 3, wr = w4 >> 1, wc = w4 & 1;
    const int total = n_row_blocks * n_col_blocks;
    int logical = blockIdx.x;
    if (total % 8 == 0) logical = (blockIdx.x % 8) * (total / 8) + blockIdx.x / 8;
    const int cb = row_panels ? logical % n_col_blocks : logical / n_row_blocks;
    const int rb = row_panels ? logical / n_col_blocks : logical % n_row_blocks;
    const int row0 = rb * kBM, col0 = cb * kBN;
    f32x16 acc[2][3];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 3; ++u)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[t][u][q] = 0.f;

    // staging: A is FP32, 1024 chunks of 4 values (two per thread), split
    // into hi/lo BF16 on the way to LDS; B 768 chunks of 8 BF16 per part
    // (threads 0..255 take a second one); three stages in flight
    typedef __attribute__((ext_vector_type(4))) uint32_t u32x4v;
    typedef __attribute__((ext_vector_type(4))) float f32x4v;
    typedef __attribute__((ext_vector_type(4))) __bf16 bf16x4;
    f32x4v ga[3][2];
    u32x4v gb[3][2][2];
    const bool two_b = tid < 256;
    const int n_stages = K / kBK;
    auto load_stage = [&](auto set_c, int st) {
        constexpr int R = decltype(set_c)::value;
        if (st >= n_stages) return;
        const int k0 = st * kBK;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int ch = tid + 512 * i, row = ch >> 3, kc = ch & 7;
            ga[R][i] = *reinterpret_cast<const f32x4v *>(a + (size_t)(row0 + row) * K + k0 + 4 * kc);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (i == 1 && !two_b) break;
            const int ch = tid + 512 * i, row = ch >> 2, kc = ch & 3;
            const size_t off = (size_t)(col0 + row) * K + k0 + 8 * kc;
            gb[R][0][i] = *reinterpret_cast<const u32x4v *>(b_hi + off);
            gb[R][1][i] = *reinterpret_cast<const u32x4v *>(b_lo + off);
        }
    };
    auto store_stage = [&](auto set_c, auto buf_c) {
        constexpr int R = decltype(set_c)::value, B = decltype(buf_c)::value;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int ch = tid + 512 * i, row = ch >> 3, kc = ch & 7;
            bf16x4 hi, lo;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float x = ga[R][i][q];
                const __bf16 h = (__bf16)x;
                hi[q] = h;
                lo[q] = (__bf16)(x - (float)h);
            }
            *reinterpret_cast<bf16x4 *>(sa(B, 0) + row * kLdsRow + 4 * kc) = hi;
            *reinterpret_cast<bf16x4 *>(sa(B, 1) + row * kLdsRow + 4 * kc) = lo;
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (i == 1 && !two_b) break;
            const int ch = tid + 512 * i, row = ch >> 2, kc = ch & 3;
#pragma unroll
            for (int h = 0; h < 2; ++h)
                *reinterpret_cast<u32x4v *>(sb(B, h) + row * kLdsRow + 8 * kc) = gb[R][h][i];
        }
    };
    auto compute = [&](auto buf_c) {
        constexpr int B = decltype(buf_c)::value;
        const int kq = 16 * kh + 8 * (l >> 5);
        bf16x8 xa[2][2], xb[2][3];
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                xa[h][t] = *reinterpret_cast<const bf16x8 *>(
                    sa(B, h) + (64 * wr + 32 * t + (l & 31)) * kLdsRow + kq);
#pragma unroll
        for (int u = 0; u < 3; ++u)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                xb[h][u] = *reinterpret_cast<const bf16x8 *>(
                    sb(B, h) + (96 * wc + 32 * u + (l & 31)) * kLdsRow + kq);
#pragma unroll
        for (int term = 0; term < 3; ++term)
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 3; ++u)
                    acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                        xa[term == 2 ? 1 : 0][t], xb[term == 1 ? 1 : 0][u], acc[t][u], 0, 0, 0);
    };
    auto step = [&](auto s_c, int st) {
        constexpr int S6 = decltype(s_c)::value;
        compute(std::integral_constant<int, S6 % 2>{});
        if (st + 1 < n_stages)
            store_stage(std::integral_constant<int, (S6 + 1) % 3>{},
                        std::integral_constant<int, (S6 + 1) % 2>{});
        __syncthreads();
        load_stage(std::integral_constant<int, (S6 + 1) % 3>{}, st + 4);
    };
    load_stage(std::integral_constant<int, 0>{}, 0);
    store_stage(std::integral_constant<int, 0>{}, std::integral_constant<int, 0>{});
    __syncthreads();
    load_stage(std::integral_constant<int, 1>{}, 1);
    load_stage(std::integral_constant<int, 2>{}, 2);
    load_stage(std::integral_constant<int, 0>{}, 3);
    for (int st = 0; st < n_stages; st += 6) {
        step(std::integral_constant<int, 0>{}, st);
        step(std::integral_constant<int, 1>{}, st + 1);
        step(std::integral_constant<int, 2>{}, st + 2);
        step(std::integral_constant<int, 3>{}, st + 3);
        step(std::integral_constant<int, 4>{}, st + 4);
        step(std::integral_constant<int, 5>{}, st + 5);
    }
    // the loop ends on a barrier: the staging LDS is free for the reduction
    // of the two k halves (4 tiles x 6144 floats = 96 KB)
    float *red = reinterpret_cast<float *>(lds);
    static_assert(sizeof(lds) >= 4 * 6144 * sizeof(float), "reduction buffer");
    if constexpr (kEpi == 0) {
        float *mine = red + (size_t)w4 * 6144;
        if (kh == 1) {
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 3; ++u)
#pragma unroll
                    for (int q = 0; q < 16; ++q) mine[((t * 3 + u) * 16 + q) * 64 + l] = acc[t][u][q];
        }
        __syncthreads();
        if (kh == 0) {
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 3; ++u) {
#pragma unroll
                    for (int q = 0; q < 16; ++q) acc[t][u][q] += mine[((t * 3 + u) * 16 + q) * 64 + l];
                    store_tile32(c, ldc, row0 + 64 * wr + 32 * t, col0 + 96 * wc + 32 * u, acc[t][u]);
                }
        }
    } else {
        // each half sends the registers the other finishes.  The stage's
        // loads of a lane's 8 elements per row tile (stage 1: state and
        // Philox keys; stage 2: state, predictor and the increments stage 1
        // drew) are issued together, the first tile's before the exchange
        // barrier.
        float *send = red + (size_t)(2 * w4 + kh) * 3072;
        const float *recv = red + (size_t)(2 * w4 + 1 - kh) * 3072;
        auto send_half = [&](auto qs_c) {
            constexpr int QS = decltype(qs_c)::value;
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 3; ++u)
#pragma unroll
                    for (int qq = 0; qq < 8; ++qq)
                        send[((t * 3 + u) * 8 + qq) * 64 + l] = acc[t][u][QS + qq];
        };
        const int gcol = col0 + 96 * wc;             // x column of this wave's 32 particles
        const int p = 32 * (gcol / 96) + (l & 31);   // particle of this lane's columns
        const ReplicaParticle cp = replica_particle(ra, p < ra.n ? p : ra.n - 1);
        const int qm = 8 * kh;                       // registers this half finishes
        struct Batch {
            float m[8][3], t[8][3], z[8][3];
            uint32_t k0[8], k1[8];
        };
        auto row_of = [&](int t, int qq) {
            const int q = qm + qq;
            return row0 + 64 * wr + 32 * t + 8 * (q >> 2) + 4 * (l >> 5) + (q & 3);
        };
        auto load_batch = [&](int t, Batch &bt) {
#pragma unroll
            for (int qq = 0; qq < 8; ++qq) {
                const int r = row_of(t, qq);
                const int rr = r < ra.reps ? r : 0;      // rows past the last replica: row 0
                const size_t ox = (size_t)rr * ra.ld + gcol + (l & 31);
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    bt.m[qq][u] = ra.m[ox + 32 * u];
                    if constexpr (kEpi == 2) {
                        bt.t[qq][u] = ra.mt[ox + 32 * u];
                        bt.z[qq][u] = ra.nz[ox + 32 * u];
                    }
                }
                if constexpr (kEpi == 1) {
                    bt.k0[qq] = ra.key0[rr];
                    bt.k1[qq] = ra.key1[rr];
                }
            }
        };
        double ex = 0.0, ey = 0.0, ez = 0.0;
        auto finish_batch = [&](auto qm_c, auto t_c, const Batch &bt) {
            constexpr int QM = decltype(qm_c)::value, t = decltype(t_c)::value;
#pragma unroll
            for (int qq = 0; qq < 8; ++qq) {
                float h[3];
#pragma unroll
                for (int u = 0; u < 3; ++u)
                    h[u] = acc[t][u][QM + qq] + recv[((t * 3 + u) * 8 + qq) * 64 + l];
                const int r = row_of(t, qq);
                const bool live = r < ra.reps && p < ra.n;
                if (!live) continue;
                const size_t ox = (size_t)r * ra.ld + gcol + (l & 31);
                if constexpr (kEpi == 1) {
                    float sx, sy, sz, tx, ty, tz;
                    replica_normals(bt.k0[qq], bt.k1[qq], ra.tag, p, n_step, cp.sig, sx, sy, sz);
                    replica_stage1_math(ra, cp, sx, sy, sz, hstage, h[0], h[1], h[2], bt.m[qq][0],
                                        bt.m[qq][1], bt.m[qq][2], tx, ty, tz);
                    ra.mt[ox] = tx;
                    ra.mt[ox + 32] = ty;
                    ra.mt[ox + 64] = tz;
                    ra.nz[ox] = sx;
                    ra.nz[ox + 32] = sy;
                    ra.nz[ox + 64] = sz;
                } else {
                    float mx, my, mz;
                    replica_stage2_math(ra, cp, bt.z[qq][0], bt.z[qq][1], bt.z[qq][2], hstage, h[0],
                                        h[1], h[2], bt.t[qq][0], bt.t[qq][1], bt.t[qq][2],
                                        bt.m[qq][0], bt.m[qq][1], bt.m[qq][2], mx, my, mz);
                    replica_store_state(ra, r, p, ox, mx, my, mz);
                    if (n_rec > 0) {
                        const double w = ra.msw[r];
                        ex += (double)mx * w;
                        ey += (double)my * w;
                        ez += (double)mz * w;
                        if (ra.samples != nullptr) {
                            const size_t P = (size_t)ra.n * ra.reps, i = (size_t)r * ra.n + p;
                            for (int k = 0; k < n_rec; ++k) {
                                float *base = ra.samples + (size_t)(s_first + k) * 3 * P;
                                base[i] = mx;
                                base[P + i] = my;
                                base[2 * P + i] = mz;
                            }
                        }
                    }
                }
            }
        };
        Batch b0, b1;
        load_batch(0, b0);
        if (kh == 0) send_half(std::integral_constant<int, 8>{});
        else send_half(std::integral_constant<int, 0>{});
        __syncthreads();
        // stage 1 holds both tiles' loads in flight; stage 2 loads twice as
        // much per tile and issues the second tile's after the first
        if constexpr (kEpi == 1) load_batch(1, b1);
        using I0 = std::integral_constant<int, 0>;
        using I1 = std::integral_constant<int, 1>;
        using I8 = std::integral_constant<int, 8>;
        if (kh == 0) finish_batch(I0{}, I0{}, b0);
        else finish_batch(I8{}, I0{}, b0);
        if constexpr (kEpi == 2) load_batch(1, b1);
        if (kh == 0) finish_batch(I0{}, I1{}, b1);
        else finish_batch(I8{}, I1{}, b1);
        if constexpr (kEpi == 2) {
            if (n_rec > 0 && ra.ens_partial != nullptr) {
                wave_sum3<float>(ex, ey, ez);
                if (l == 0) {
                    const int wid = blockIdx.x * 8 + wave;
                    for (int k = 0; k < n_rec; ++k) {
                        double *dst = ra.ens_partial + (size_t)(s_first + k) * 3 * (size_t)ra.n_waves;
                        dst[wid] = ex;
                        dst[(size_t)ra.n_waves + wid] = ey;
                        dst[2 * (size_t)ra.n_waves + wid] = ez;
                    }
                }
            }
        }
    }
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

// The stages as elementwise kernels around an unfused product (FP32 MFMA
// product): thread = (replica r, particle p), p fastest.
static __global__ __launch_bounds__(256) void replica_stage1_kernel(ReplicaArgs a, int64_t n_step, float h1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n * a.reps) return;
    const int r = i / a.n, p = i % a.n;
    const size_t ox = (size_t)r * a.ld + replica_col(p, 0);
    const ReplicaParticle c = replica_particle(a, p);
    float sx, sy, sz, tx, ty, tz;
    replica_normals(a.key0[r], a.key1[r], a.tag, p, n_step, c.sig, sx, sy, sz);
    replica_stage1_math(a, c, sx, sy, sz, h1, a.h[ox], a.h[ox + 32], a.h[ox + 64], a.m[ox],
                        a.m[ox + 32], a.m[ox + 64], tx, ty, tz);
    a.mt[ox] = tx;
    a.mt[ox + 32] = ty;
    a.mt[ox + 64] = tz;
}

// stage 2 and the samples s_first .. s_first + n_rec - 1 (all recorded after
// this step)
static __global__ __launch_bounds__(256) void replica_stage2_kernel(ReplicaArgs a, int64_t n_step, float h2,
                                                             int64_t s_first, int n_rec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < a.n * a.reps;
    const int q = live ? i : a.n * a.reps - 1;
    const int r = q / a.n, p = q % a.n;
    const size_t ox = (size_t)r * a.ld + replica_col(p, 0);
    const ReplicaParticle c = replica_particle(a, p);
    float sx, sy, sz, mx, my, mz;
    replica_normals(a.key0[r], a.key1[r], a.tag, p, n_step, c.sig, sx, sy, sz);
    replica_stage2_math(a, c, sx, sy, sz, h2, a.h[ox], a.h[ox + 32], a.h[ox + 64], a.mt[ox],
                        a.mt[ox + 32], a.mt[ox + 64], a.m[ox], a.m[ox + 32], a.m[ox + 64], mx, my,
                        mz);
    if (live) replica_store_state(a, r, p, ox, mx, my, mz);
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
