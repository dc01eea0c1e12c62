// Learned probe decoder: a small space-time convolutional network that reads
// a trajectory's measurement record (int8 [H=time][W=sites], -1 unmeasured,
// 0, 1) and predicts the probe qubit's conditional Bloch vector, trained on
// records and exact labels the simulator generates.  No counterpart in the
// reference.
//
// Network: input channels (record == 0, record == 1, x == W/2 (the probe's
// site), (y + 1) / H (time fraction)); four 3x3 convolutions
// (zero padding in time, periodic in sites: the chain is a ring) to 64
// channels with bias and ReLU; mean pooling over the H*W positions; 64 -> 64
// (ReLU) -> 3.  Loss: batch mean of h(S(|out|) - S(|y|)) + (|y| - out.y/|out|),
// h the smooth-L1 function of width delta (d^2/(2 delta) inside, |d| - delta/2
// outside; d^2 when delta = 0; default 0.05),
// with |v| = sqrt(sum v^2 + 1e-12) and S(r) the qubit entropy (nats) at
// Bloch length r, continued linearly past 1 - 1e-6 (no flat region above
// 1): entropy and direction are fitted separately, so a direction the record
// leaves open does not shrink the predicted length.  Adam (lr given per call, betas 0.9/0.999,
// eps 1e-8) on FP32 master weights.
//
// Precision: weights and activations are bfloat16 where they feed the
// matrix cores; every product accumulates in FP32.  The convolutions are
// implicit GEMMs on v_mfma_f32_16x16x32_bf16 (64x64 output tile per
// 256-thread workgroup, K in steps of 32 staged through LDS):
//   forward      [B*H*W positions] x [64 out]   over K = 9*CI (tap, ci)
//   input grad   [B*H*W positions] x [64 in]    over K = 9*64 (tap, co)
//   weight grad  [9*CI (tap, ci)]  x [64 out]   over the positions, split
//                into chunks whose FP32 partials are summed in a fixed order.
// Gradients wrt activations stay FP32 in HBM and are rounded to bfloat16
// where they enter a matrix-core product; bias gradients sum the FP32 values.
// The small pooling/MLP/loss/Adam steps are FP32 VALU kernels with fixed
// summation orders, so a training step is deterministic.
#include "mipt_hip.h"
#include "common.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

namespace mipt_hip
{
namespace
{

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int CH = 64;        // channels of every hidden layer
constexpr int NT = 256;       // threads per GEMM workgroup (4 waves)
constexpr int TM = 64;        // output tile rows
constexpr int TK = 32;        // K step (one MFMA 16x16x32)
constexpr int LP = TK + 8;    // LDS row stride of a K step (bf16), 80 B
constexpr int CP = CH + 8;    // LDS row stride of the epilogue tile (bf16), 144 B
constexpr int WG_BLOCKS = 2048;   // first-layer weight-gradient grid: 8 blocks on each of 256 CUs
constexpr int WG3_BLOCKS = 1024;  // grid of the three-tap weight gradient (4 blocks per CU)
constexpr float NORM_EPS = 1e-12f;
constexpr float MU = 1.0f;            // weight of the direction term
constexpr float R_MAX = 1.0f - 1e-6f;  // Bloch length cap of the entropy term
constexpr int BROWS = 256;             // positions per bias-gradient partial

__device__ inline void position(int m, int HW, int W, int &b, int &y, int &x)
{
    b = m / HW;
    const int r = m - b * HW;
    y = r / W;
    x = r - y * W;
}

__device__ inline floatx4 mfma(const bf16x8 &a, const bf16x8 &b, const floatx4 &c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline bf16x8 zero8()
{
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; ++i) z[i] = static_cast<bf16>(0.f);
    return z;
}

constexpr int CI0 = 4;  // input channels of the first convolution

// Input channel c at (y, x) of sample s: record == 0, record == 1, the
// probe's column, time fraction; zero outside the time range.
__device__ inline float input_channel(const std::int8_t *rec, int s, int yy, int xx, int H, int W, int c)
{
    if (yy < 0 || yy >= H) return 0.f;
    if (c == 2) return xx == W / 2 ? 1.f : 0.f;
    if (c == 3) return static_cast<float>(yy + 1) / static_cast<float>(H);
    return (rec[(static_cast<std::size_t>(s) * H + yy) * W + xx] == c) ? 1.f : 0.f;
}

// One K step of a 64x64 tile: waves own 32x32 quadrants as 2x2 MFMA tiles.
// A is [64 rows][K step] and B is [64 cols][K step] in LDS (K contiguous).
__device__ inline void tile_step(const bf16 *As, const bf16 *Bs, floatx4 acc[2][2], int lane, int wave)
{
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int r = lane & 15, kq = (lane >> 4) * 8;
    const bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(As + (wr + r) * LP + kq);
    const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(As + (wr + 16 + r) * LP + kq);
    const bf16x8 b0 = *reinterpret_cast<const bf16x8 *>(Bs + (wc + r) * LP + kq);
    const bf16x8 b1 = *reinterpret_cast<const bf16x8 *>(Bs + (wc + 16 + r) * LP + kq);
    acc[0][0] = mfma(a0, b0, acc[0][0]);
    acc[0][1] = mfma(a0, b1, acc[0][1]);
    acc[1][0] = mfma(a1, b0, acc[1][0]);
    acc[1][1] = mfma(a1, b1, acc[1][1]);
}

// D of a 16x16 MFMA tile: lane holds rows 4*(lane/16)+e of column lane%16.
template <class F>
__device__ inline void for_tile_outputs(const floatx4 acc[2][2], int lane, int wave, F &&f)
{
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) f(wr + 16 * i + 4 * (lane >> 4) + e, wc + 16 * j + (lane & 15), acc[i][j][e]);
}

__device__ inline float qubit_entropy(float r)
{
    const float c = fminf(r, R_MAX);
    const float a = 0.5f * (1.f + c), b = 0.5f * (1.f - c);
    return -a * logf(a) - b * logf(b) - atanhf(R_MAX) * fmaxf(r - R_MAX, 0.f);
}

// ---------------------------------------------------------------- forward
// out[m][co] = relu(bias[co] + sum_{tap,ci} wf[co][tap*CI+ci] * in[m + off(tap)][ci])
// FIRST: the input is the record (4 channels computed on the fly, K = 36
// padded to 64), rows of the batch gathered through idx.
template <bool FIRST>
__global__ void __launch_bounds__(NT) conv_fwd_kernel(const bf16 *__restrict__ in, const std::int8_t *__restrict__ rec,
                                                      const int *__restrict__ idx, const bf16 *__restrict__ wf,
                                                      const float *__restrict__ bias, bf16 *__restrict__ out, int M,
                                                      int H, int W, int KP)
{
    __shared__ __attribute__((aligned(16))) bf16 sm[2 * TM * LP];
    bf16 *As = sm, *Bs = sm + TM * LP;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int HW = H * W;
    const int m0 = blockIdx.x * TM;
    const int r = tid >> 2, ks = (tid & 3) * 8;
    const int m = m0 + r;
    int b, y, x;
    position(m, HW, W, b, y, x);
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < KP; k0 += TK)
    {
        bf16x8 av = zero8();
        if (FIRST)
        {
            const int s = (m < M) ? (idx ? idx[b] : b) : 0;
#pragma unroll
            for (int i = 0; i < 8; ++i)
            {
                const int k = k0 + ks + i, tap = k / CI0, c = k % CI0;
                float v = 0.f;
                if (k < 9 * CI0 && m < M)
                    v = input_channel(rec, s, y + tap / 3 - 1, (x + tap % 3 - 1 + W) % W, H, W, c);
                av[i] = static_cast<bf16>(v);
            }
        }
        else
        {
            const int tap = k0 >> 6, ci0 = (k0 & 63) + ks;
            const int yy = y + tap / 3 - 1, xx = (x + tap % 3 - 1 + W) % W;
            if (m < M && yy >= 0 && yy < H)
                av = *reinterpret_cast<const bf16x8 *>(in + ((static_cast<std::size_t>(b) * H + yy) * W + xx) * CH + ci0);
        }
        *reinterpret_cast<bf16x8 *>(As + r * LP + ks) = av;
        *reinterpret_cast<bf16x8 *>(Bs + r * LP + ks) =
            *reinterpret_cast<const bf16x8 *>(wf + static_cast<std::size_t>(r) * KP + k0 + ks);
        __syncthreads();
        tile_step(As, Bs, acc, lane, wave);
        __syncthreads();
    }
    // epilogue: bias, ReLU, bf16, through LDS for 16-byte row stores
    bf16 *Cs = sm;
    for_tile_outputs(acc, lane, wave, [&](int row, int col, float v) {
        v += bias[col];
        Cs[row * CP + col] = static_cast<bf16>(v > 0.f ? v : 0.f);
    });
    __syncthreads();
    for (int sidx = tid; sidx < TM * (CH / 8); sidx += NT)
    {
        const int row = sidx >> 3, seg = (sidx & 7) * 8;
        if (m0 + row < M)
            *reinterpret_cast<bf16x8 *>(out + static_cast<std::size_t>(m0 + row) * CH + seg) =
                *reinterpret_cast<const bf16x8 *>(Cs + row * CP + seg);
    }
}

// ------------------------------------------------------------- input grad
// din[q][ci] = sum_{tap,co} wd[ci][tap*64+co] * dz[q - off(tap)][co] (zero
// where q - off(tap) leaves the time range; periodic in sites).
__global__ void __launch_bounds__(NT) conv_bwd_data_kernel(const bf16 *__restrict__ dz, const bf16 *__restrict__ wd,
                                                           float *__restrict__ din, int M, int H, int W)
{
    __shared__ __attribute__((aligned(16))) bf16 sm[2 * TM * LP];
    bf16 *As = sm, *Bs = sm + TM * LP;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int HW = H * W;
    const int m0 = blockIdx.x * TM;
    const int r = tid >> 2, ks = (tid & 3) * 8;
    const int m = m0 + r;
    int b, y, x;
    position(m, HW, W, b, y, x);
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    constexpr int K = 9 * CH;
    for (int k0 = 0; k0 < K; k0 += TK)
    {
        const int tap = k0 >> 6, co0 = (k0 & 63) + ks;
        const int yy = y - (tap / 3 - 1), xx = (x - (tap % 3 - 1) + W) % W;
        bf16x8 av = zero8();
        if (m < M && yy >= 0 && yy < H)
            av = *reinterpret_cast<const bf16x8 *>(dz + ((static_cast<std::size_t>(b) * H + yy) * W + xx) * CH + co0);
        *reinterpret_cast<bf16x8 *>(As + r * LP + ks) = av;
        *reinterpret_cast<bf16x8 *>(Bs + r * LP + ks) =
            *reinterpret_cast<const bf16x8 *>(wd + static_cast<std::size_t>(r) * K + k0 + ks);
        __syncthreads();
        tile_step(As, Bs, acc, lane, wave);
        __syncthreads();
    }
    for_tile_outputs(acc, lane, wave, [&](int row, int col, float v) {
        if (m0 + row < M) din[static_cast<std::size_t>(m0 + row) * CH + col] = v;
    });
}

// ------------------------------------------------------------ weight grad
// partial[chunk][k*64 + co] = sum over the chunk's positions m of
// in_shift(tap)[m][ci] * dz[m][co], k = tap*CI + ci; grid (k tiles, chunks).
template <bool FIRST>
__global__ void __launch_bounds__(NT) conv_bwd_weight_kernel(const bf16 *__restrict__ in,
                                                             const std::int8_t *__restrict__ rec,
                                                             const int *__restrict__ idx, const bf16 *__restrict__ dz,
                                                             float *__restrict__ partial,
                                                             int M, int H, int W, int krows, int chunk_len)
{
    __shared__ __attribute__((aligned(16))) bf16 sm[2 * TM * LP];
    bf16 *As = sm, *Bs = sm + TM * LP;  // As[k][m], Bs[co][m]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int HW = H * W;
    const int kt = blockIdx.x, chunk = blockIdx.y;
    const int mm = tid >> 3, seg = (tid & 7) * 8;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int mbeg = chunk * chunk_len, mend = std::min(M, mbeg + chunk_len);
    // Staging of dz and of the shifted bf16 input: thread (channel pair cp,
    // position quad mq) loads channels 2cp, 2cp+1 of positions ms+4mq..+3
    // and writes each channel's four values as one 8-byte LDS store.
    const int cp = tid & 31, mq = tid >> 5;
    for (int ms = mbeg; ms < mend; ms += TK)
    {
        std::uint32_t va[4] = {0u, 0u, 0u, 0u}, vb[4] = {0u, 0u, 0u, 0u};
        {
            int b, y, x;
            position(ms + 4 * mq, HW, W, b, y, x);
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const int m = ms + 4 * mq + j;
                if (m < mend)
                {
                    vb[j] = *reinterpret_cast<const std::uint32_t *>(dz + static_cast<std::size_t>(m) * CH + 2 * cp);
                    if (!FIRST)
                    {
                        const int yy = y + kt / 3 - 1;
                        int xx = x + kt % 3 - 1;
                        xx += xx < 0 ? W : 0;
                        xx -= xx >= W ? W : 0;
                        if (yy >= 0 && yy < H)
                            va[j] = *reinterpret_cast<const std::uint32_t *>(
                                in + ((static_cast<std::size_t>(b) * H + yy) * W + xx) * CH + 2 * cp);
                    }
                }
                if (++x == W)
                {
                    x = 0;
                    if (++y == H)
                    {
                        y = 0;
                        ++b;
                    }
                }
            }
        }
        auto store_pair = [&](bf16 *S, const std::uint32_t v[4]) {
            const uint2 lo = {(v[0] & 0xffffu) | (v[1] << 16), (v[2] & 0xffffu) | (v[3] << 16)};
            const uint2 hi = {(v[0] >> 16) | (v[1] & 0xffff0000u), (v[2] >> 16) | (v[3] & 0xffff0000u)};
            *reinterpret_cast<uint2 *>(S + (2 * cp) * LP + 4 * mq) = lo;
            *reinterpret_cast<uint2 *>(S + (2 * cp + 1) * LP + 4 * mq) = hi;
        };
        store_pair(Bs, vb);
        // A^T: rows k = kt*64 + ci, column m
        if (FIRST)
        {
            const int m = ms + mm;
            int b, y, x;
            position(m, HW, W, b, y, x);
            const bool valid = m < mend;
            const int s = valid ? (idx ? idx[b] : b) : 0;
#pragma unroll
            for (int i = 0; i < 8; ++i)
            {
                const int k = seg + i, tap = k / CI0, c = k % CI0;
                float v = 0.f;
                if (valid && k < 9 * CI0) v = input_channel(rec, s, y + tap / 3 - 1, (x + tap % 3 - 1 + W) % W, H, W, c);
                As[(seg + i) * LP + mm] = static_cast<bf16>(v);
            }
        }
        else
            store_pair(As, va);
        __syncthreads();
        tile_step(As, Bs, acc, lane, wave);
        __syncthreads();
    }
    float *dst = partial + static_cast<std::size_t>(chunk) * krows * CH;
    for_tile_outputs(acc, lane, wave, [&](int row, int col, float v) {
        const int k = kt * TM + row;
        if (k < krows) dst[static_cast<std::size_t>(k) * CH + col] = v;
    });
}

// Layers 1..3 (64 input channels): one block per (row offset dy = g - 1,
// chunk) computes the three taps (g, dx = -1, 0, 1), k = (3g + dx + 1)*64 + ci,
// so the dz tile is staged once per K step for three MFMA tiles; the three
// shifted inputs are neighbouring positions of the same row.  Staging as in
// conv_bwd_weight_kernel (8-byte LDS stores of four positions).
__device__ inline void store_quad(bf16 *S, const std::uint32_t v[4], int cp, int mq)
{
    const uint2 lo = {(v[0] & 0xffffu) | (v[1] << 16), (v[2] & 0xffffu) | (v[3] << 16)};
    const uint2 hi = {(v[0] >> 16) | (v[1] & 0xffff0000u), (v[2] >> 16) | (v[3] & 0xffff0000u)};
    *reinterpret_cast<uint2 *>(S + (2 * cp) * LP + 4 * mq) = lo;
    *reinterpret_cast<uint2 *>(S + (2 * cp + 1) * LP + 4 * mq) = hi;
}

__global__ void __launch_bounds__(NT) conv_bwd_weight3_kernel(const bf16 *__restrict__ in, const bf16 *__restrict__ dz,
                                                              float *__restrict__ partial, int M, int H, int W,
                                                              int chunk_len)
{
    __shared__ __attribute__((aligned(16))) bf16 sm[4 * TM * LP];
    bf16 *Bs = sm + 3 * TM * LP;  // As[tt] = sm + tt*TM*LP: [ci][m]; Bs: [co][m]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int HW = H * W;
    const int g = blockIdx.x, chunk = blockIdx.y;
    floatx4 acc[3][2][2];
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[t][i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int mbeg = chunk * chunk_len, mend = std::min(M, mbeg + chunk_len);
    const int cp = tid & 31, mq = tid >> 5;
    for (int ms = mbeg; ms < mend; ms += TK)
    {
        std::uint32_t va[3][4], vb[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            vb[j] = 0u;
#pragma unroll
            for (int t = 0; t < 3; ++t) va[t][j] = 0u;
        }
        {
            int b, y, x;
            position(ms + 4 * mq, HW, W, b, y, x);
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const int m = ms + 4 * mq + j;
                if (m < mend)
                {
                    vb[j] = *reinterpret_cast<const std::uint32_t *>(dz + static_cast<std::size_t>(m) * CH + 2 * cp);
                    const int yy = y + g - 1;
                    if (yy >= 0 && yy < H)
                    {
                        const std::size_t row = (static_cast<std::size_t>(b) * H + yy) * W;
#pragma unroll
                        for (int t = 0; t < 3; ++t)
                        {
                            int xx = x + t - 1;
                            xx += xx < 0 ? W : 0;
                            xx -= xx >= W ? W : 0;
                            va[t][j] = *reinterpret_cast<const std::uint32_t *>(in + (row + xx) * CH + 2 * cp);
                        }
                    }
                }
                if (++x == W)
                {
                    x = 0;
                    if (++y == H)
                    {
                        y = 0;
                        ++b;
                    }
                }
            }
        }
        store_quad(Bs, vb, cp, mq);
#pragma unroll
        for (int t = 0; t < 3; ++t) store_quad(sm + t * TM * LP, va[t], cp, mq);
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 3; ++t) tile_step(sm + t * TM * LP, Bs, acc[t], lane, wave);
        __syncthreads();
    }
    float *dst = partial + static_cast<std::size_t>(chunk) * 9 * CH * CH;
#pragma unroll
    for (int t = 0; t < 3; ++t)
        for_tile_outputs(acc[t], lane, wave, [&](int row, int col, float v) {
            dst[static_cast<std::size_t>((3 * g + t) * TM + row) * CH + col] = v;
        });
}

// grad[e] = sum over chunks of partial[chunk][e].  A block covers 64
// consecutive elements; wave w takes chunks w, w+4, w+8, ... (two running
// sums), and the four wave sums are added in a fixed order.
__global__ void __launch_bounds__(256) sum_partials_kernel(const float *__restrict__ partial, int nchunks, int n,
                                                           float *__restrict__ grad)
{
    __shared__ float red[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = blockIdx.x * 64 + lane;
    float s0 = 0.f, s1 = 0.f;
    if (e < n)
    {
        int c = wave;
#pragma unroll 4
        for (; c + 4 < nchunks; c += 8)
        {
            s0 += partial[static_cast<std::size_t>(c) * n + e];
            s1 += partial[static_cast<std::size_t>(c + 4) * n + e];
        }
        if (c < nchunks) s0 += partial[static_cast<std::size_t>(c) * n + e];
    }
    red[wave][lane] = s0 + s1;
    __syncthreads();
    if (wave == 0 && e < n) grad[e] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// out[c] = sum over rows r of part[r][c]: one block per column, rows split
// over 256 threads (r = t, t+256, ...) and added in a fixed tree order
__global__ void __launch_bounds__(256) sum_columns_kernel(const float *__restrict__ part, int nrows, int ncols,
                                                          float *__restrict__ out)
{
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < nrows; r += 256) s += part[static_cast<std::size_t>(r) * ncols + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1)
    {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// dz[m][c] = bf16(dact[m][c] if act[m][c] > 0 else 0), the gradient at the
// convolution's output that both backward GEMMs read, and bpart[block][c] =
// the sum of the unrounded values over the block's BROWS positions.  Thread
// (row group tid/8, channels 8*(tid%8)...+7) takes rows g, g+32, ...; the 32
// group sums are added in order.
__global__ void __launch_bounds__(256) dz_bias_kernel(const float *__restrict__ dact, const bf16 *__restrict__ act,
                                                      int M, bf16 *__restrict__ dz, float *__restrict__ bpart)
{
    __shared__ float red[32][CH + 1];
    const int grp = threadIdx.x >> 3, seg = (threadIdx.x & 7) * 8;
    const int r0 = blockIdx.x * BROWS;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int r = grp; r < BROWS; r += 32)
    {
        const int m = r0 + r;
        if (m >= M) break;
        const std::size_t o = static_cast<std::size_t>(m) * CH + seg;
        const float4 g0 = *reinterpret_cast<const float4 *>(dact + o);
        const float4 g1 = *reinterpret_cast<const float4 *>(dact + o + 4);
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(act + o);
        const float g[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
        bf16x8 z;
#pragma unroll
        for (int i = 0; i < 8; ++i)
        {
            const float v = static_cast<float>(a[i]) > 0.f ? g[i] : 0.f;
            s[i] += v;
            z[i] = static_cast<bf16>(v);
        }
        *reinterpret_cast<bf16x8 *>(dz + o) = z;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) red[grp][seg + i] = s[i];
    __syncthreads();
    if (threadIdx.x < CH)
    {
        float t = 0.f;
        for (int k = 0; k < 32; ++k) t += red[k][threadIdx.x];
        bpart[blockIdx.x * CH + threadIdx.x] = t;
    }
}

// ------------------------------------------------------ pooling, MLP, loss
// pooled[b][c] = mean over positions; 256 threads = 4 interleaved position sets
__global__ void __launch_bounds__(256) pool_fwd_kernel(const bf16 *__restrict__ act, int HW,
                                                       float *__restrict__ pooled)
{
    __shared__ float red[4][CH];
    const int b = blockIdx.x, c = threadIdx.x & (CH - 1), q = threadIdx.x >> 6;
    const bf16 *p = act + static_cast<std::size_t>(b) * HW * CH + c;
    float s = 0.f;
    for (int i = q; i < HW; i += 4) s += static_cast<float>(p[static_cast<std::size_t>(i) * CH]);
    red[q][c] = s;
    __syncthreads();
    if (q == 0) pooled[b * CH + c] = ((red[0][c] + red[1][c]) + (red[2][c] + red[3][c])) / static_cast<float>(HW);
}

// h = relu(pooled @ w1 + b1); out = h @ w2 + b2 (bf16 weights, FP32 math)
__global__ void mlp_fwd_kernel(const float *__restrict__ pooled, const bf16 *__restrict__ w1, const float *__restrict__ b1,
                               const bf16 *__restrict__ w2, const float *__restrict__ b2, float *__restrict__ h,
                               float *__restrict__ out)
{
    __shared__ float sp[CH], sh[CH];
    const int b = blockIdx.x, j = threadIdx.x;
    sp[j] = pooled[b * CH + j];
    __syncthreads();
    float a = b1[j];
    for (int c = 0; c < CH; ++c) a += static_cast<float>(w1[c * CH + j]) * sp[c];
    a = a > 0.f ? a : 0.f;
    h[b * CH + j] = a;
    sh[j] = a;
    __syncthreads();
    if (j < 3)
    {
        float o = b2[j];
        for (int q = 0; q < CH; ++q) o += static_cast<float>(w2[q * 3 + j]) * sh[q];
        out[b * 3 + j] = o;
    }
}

// per sample: loss term, dout, dh, dpooled (batch size nb for the 2/B factor)
__global__ void loss_bwd_kernel(const float *__restrict__ out, const float *__restrict__ labels,
                                const int *__restrict__ idx, const float *__restrict__ h, const bf16 *__restrict__ w1,
                                const bf16 *__restrict__ w2, int nb, float huber, float *__restrict__ loss,
                                float *__restrict__ dout, float *__restrict__ dh, float *__restrict__ dpooled)
{
    __shared__ float so[3], sdh[CH];
    const int b = blockIdx.x, j = threadIdx.x;
    if (j == 0)
    {
        const int s = idx ? idx[b] : b;
        float o[3], yv[3], dot = 0.f, no = 0.f, ny = 0.f;
        for (int q = 0; q < 3; ++q)
        {
            o[q] = out[b * 3 + q];
            yv[q] = labels[s * 3 + q];
            dot += o[q] * yv[q];
            no += o[q] * o[q];
            ny += yv[q] * yv[q];
        }
        no = sqrtf(no + NORM_EPS);
        ny = sqrtf(ny + NORM_EPS);
        const float d = qubit_entropy(no) - qubit_entropy(ny);
        float hl, hg;  // smooth-L1 value and slope in d
        if (huber <= 0.f)
        {
            hl = d * d;
            hg = 2.f * d;
        }
        else if (fabsf(d) < huber)
        {
            hl = d * d / (2.f * huber);
            hg = d / huber;
        }
        else
        {
            hl = fabsf(d) - 0.5f * huber;
            hg = d > 0.f ? 1.f : -1.f;
        }
        loss[b] = hl + MU * (ny - dot / no);
        // dS/dr = -atanh(min(r, cap))
        const float gs = -hg * atanhf(fminf(no, R_MAX));
        const float f = 1.f / static_cast<float>(nb);
        for (int q = 0; q < 3; ++q)
        {
            so[q] = f * (gs / no * o[q] - MU * (yv[q] / no - dot / (no * no * no) * o[q]));
            dout[b * 3 + q] = so[q];
        }
    }
    __syncthreads();
    const float hj = h[b * CH + j];
    float g = 0.f;
    if (hj > 0.f)
        for (int q = 0; q < 3; ++q) g += static_cast<float>(w2[j * 3 + q]) * so[q];
    dh[b * CH + j] = g;
    sdh[j] = g;
    __syncthreads();
    float dp = 0.f;
    for (int q = 0; q < CH; ++q) dp += static_cast<float>(w1[j * CH + q]) * sdh[q];
    dpooled[b * CH + j] = dp;
}

// FC gradients, one block per parameter, the batch split over 256 threads
// and added in a fixed tree order.  Layout of g: w1[64][64], b1[64],
// w2[64][3], b2[3].
constexpr int N_FC = CH * CH + CH + CH * 3 + 3;

__global__ void __launch_bounds__(256) mlp_grad_kernel(const float *__restrict__ pooled, const float *__restrict__ h,
                                                       const float *__restrict__ dh, const float *__restrict__ dout,
                                                       int nb, float *__restrict__ g)
{
    __shared__ float red[256];
    const int e = blockIdx.x;
    constexpr int NW1 = CH * CH, NB1 = CH, NW2 = CH * 3;
    float s = 0.f;
    if (e < NW1)
    {
        const int c = e / CH, j = e % CH;
        for (int b = threadIdx.x; b < nb; b += 256) s += pooled[b * CH + c] * dh[b * CH + j];
    }
    else if (e < NW1 + NB1)
    {
        const int j = e - NW1;
        for (int b = threadIdx.x; b < nb; b += 256) s += dh[b * CH + j];
    }
    else if (e < NW1 + NB1 + NW2)
    {
        const int q = e - NW1 - NB1, j = q / 3, o = q % 3;
        for (int b = threadIdx.x; b < nb; b += 256) s += h[b * CH + j] * dout[b * 3 + o];
    }
    else
    {
        const int o = e - NW1 - NB1 - NW2;
        for (int b = threadIdx.x; b < nb; b += 256) s += dout[b * 3 + o];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1)
    {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) g[e] = red[0];
}

// d act4[m][c] = dpooled[b(m)][c] / HW
__global__ void pool_bwd_kernel(const float *__restrict__ dpooled, int M, int HW, float *__restrict__ dact)
{
    const std::size_t e = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
    if (e >= static_cast<std::size_t>(M) * CH) return;
    const int m = static_cast<int>(e / CH), c = static_cast<int>(e % CH);
    dact[e] = dpooled[(m / HW) * CH + c] / static_cast<float>(HW);
}

// --------------------------------------------------------------- optimiser
__global__ void adam_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                            const float *__restrict__ g, int n, float lr, float c1, float c2)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float ge = g[e];
    const float me = 0.9f * m[e] + 0.1f * ge;
    const float ve = 0.999f * v[e] + 0.001f * ge * ge;
    m[e] = me;
    v[e] = ve;
    w[e] -= lr * (me / c1) / (sqrtf(ve / c2) + 1e-8f);
}

// bf16 copies of the FP32 master weights in the layouts the kernels read:
// wf[co][tap*CI+ci] (forward, K padded to KP), wd[ci][tap*64+co] (input grad)
__global__ void pack_conv_kernel(const float *__restrict__ w, int CI, int KP, bf16 *__restrict__ wf,
                                 bf16 *__restrict__ wd)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= CH * KP) return;
    const int co = e / KP, k = e % KP;
    const float val = (k < 9 * CI) ? w[static_cast<std::size_t>(k) * CH + co] : 0.f;
    wf[e] = static_cast<bf16>(val);
    if (wd && k < 9 * CI)
    {
        const int tap = k / CI, ci = k % CI;
        wd[static_cast<std::size_t>(ci) * (9 * CH) + tap * CH + co] = static_cast<bf16>(val);
    }
}

__global__ void pack_plain_kernel(const float *__restrict__ w, int n, bf16 *__restrict__ out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = static_cast<bf16>(w[e]);
}

int blocks(std::size_t n, int t) { return static_cast<int>((n + t - 1) / t); }

} // namespace

// ------------------------------------------------------------------ host
struct Decoder
{
    int H = 0, W = 0, HW = 0, bmax = 0;
    hipStream_t stream = nullptr;
    // parameter layout (FP32 master), as [9][CI][64] + [64] per conv, then fc
    int ci[4] = {CI0, CH, CH, CH};
    int kp[4] = {64, 9 * CH, 9 * CH, 9 * CH};  // K = 9*CI padded to the 32-wide K step
    std::size_t off_w[4], off_b[4], off_fc = 0;
    std::size_t nparam = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
    bf16 *wf[4] = {}, *wd[4] = {}, *w1b = nullptr, *w2b = nullptr;
    bf16 *act[4] = {};
    float *dact = nullptr;  // gradient at a convolution's output (FP32)
    bf16 *dz = nullptr;     // the same through the ReLU, bf16
    float *partial = nullptr, *bpart = nullptr;
    float *pooled = nullptr, *h = nullptr, *out = nullptr, *dout = nullptr, *dh = nullptr, *dpooled = nullptr,
          *loss = nullptr;
    int *d_idx = nullptr;
    std::int8_t *data_rec = nullptr;
    float *data_lab = nullptr;
    int data_count = 0;
    long step = 0;
    float huber = 0.05f;  // smooth-L1 width of the entropy term (0: squared error)

    ~Decoder()
    {
        void *ptrs[] = {w, m, v, g, w1b, w2b, partial, bpart, pooled, h, out, dout, dh, dpooled, loss, d_idx,
                        data_rec, data_lab, dact, dz};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        for (int l = 0; l < 4; ++l)
        {
            if (wf[l]) (void)hipFree(wf[l]);
            if (wd[l]) (void)hipFree(wd[l]);
            if (act[l]) (void)hipFree(act[l]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    int init(int height, int width, int batch_max, std::uint64_t seed)
    {
        H = height;
        W = width;
        HW = H * W;
        bmax = batch_max;
        std::size_t o = 0;
        for (int l = 0; l < 4; ++l)
        {
            off_w[l] = o;
            o += static_cast<std::size_t>(9) * ci[l] * CH;
            off_b[l] = o;
            o += CH;
        }
        off_fc = o;
        o += CH * CH + CH + CH * 3 + 3;
        nparam = o;
        MIPT_HIP_CHECK(hipStreamCreate(&stream));
        MIPT_HIP_CHECK(hipMalloc(&w, nparam * 4));
        MIPT_HIP_CHECK(hipMalloc(&m, nparam * 4));
        MIPT_HIP_CHECK(hipMalloc(&v, nparam * 4));
        MIPT_HIP_CHECK(hipMalloc(&g, nparam * 4));
        MIPT_HIP_CHECK(hipMemsetAsync(m, 0, nparam * 4, stream));
        MIPT_HIP_CHECK(hipMemsetAsync(v, 0, nparam * 4, stream));
        for (int l = 0; l < 4; ++l)
        {
            MIPT_HIP_CHECK(hipMalloc(&wf[l], static_cast<std::size_t>(CH) * kp[l] * 2));
            if (l > 0) MIPT_HIP_CHECK(hipMalloc(&wd[l], static_cast<std::size_t>(CH) * 9 * CH * 2));
        }
        MIPT_HIP_CHECK(hipMalloc(&w1b, CH * CH * 2));
        MIPT_HIP_CHECK(hipMalloc(&w2b, CH * 3 * 2));
        const std::size_t M = static_cast<std::size_t>(bmax) * HW;
        for (int l = 0; l < 4; ++l) MIPT_HIP_CHECK(hipMalloc(&act[l], M * CH * 2));
        MIPT_HIP_CHECK(hipMalloc(&dact, M * CH * 4));
        MIPT_HIP_CHECK(hipMalloc(&dz, M * CH * 2));
        std::size_t psize = 0;  // at most WG_BLOCKS / ktiles chunks of krows x 64 partials per layer
        for (int l = 0; l < 4; ++l)
        {
            const int krows = 9 * ci[l];
            const int per = l == 0 ? blocks(WG_BLOCKS, blocks(krows, TM)) : blocks(WG3_BLOCKS, 3);
            psize = std::max(psize, static_cast<std::size_t>(per) * krows * CH * 4);
        }
        MIPT_HIP_CHECK(hipMalloc(&partial, psize));
        MIPT_HIP_CHECK(hipMalloc(&bpart, ((M + BROWS - 1) / BROWS) * CH * 4));
        MIPT_HIP_CHECK(hipMalloc(&pooled, bmax * CH * 4));
        MIPT_HIP_CHECK(hipMalloc(&h, bmax * CH * 4));
        MIPT_HIP_CHECK(hipMalloc(&dh, bmax * CH * 4));
        MIPT_HIP_CHECK(hipMalloc(&dpooled, bmax * CH * 4));
        MIPT_HIP_CHECK(hipMalloc(&out, bmax * 3 * 4));
        MIPT_HIP_CHECK(hipMalloc(&dout, bmax * 3 * 4));
        MIPT_HIP_CHECK(hipMalloc(&loss, bmax * 4));
        MIPT_HIP_CHECK(hipMalloc(&d_idx, bmax * 4));
        // He-normal weights, zero biases
        std::vector<float> hw(nparam, 0.f);
        std::mt19937_64 rng(seed);
        for (int l = 0; l < 4; ++l)
        {
            std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * ci[l])));
            for (std::size_t i = 0; i < static_cast<std::size_t>(9) * ci[l] * CH; ++i) hw[off_w[l] + i] = nd(rng);
        }
        std::normal_distribution<float> n1(0.f, std::sqrt(2.f / CH));
        for (int i = 0; i < CH * CH; ++i) hw[off_fc + i] = n1(rng);
        for (int i = 0; i < CH * 3; ++i) hw[off_fc + CH * CH + CH + i] = n1(rng);
        return set_params(hw.data());
    }

    int pack()
    {
        for (int l = 0; l < 4; ++l)
        {
            hipLaunchKernelGGL(pack_conv_kernel, dim3(blocks(static_cast<std::size_t>(CH) * kp[l], 256)), dim3(256), 0,
                               stream, w + off_w[l], ci[l], kp[l], wf[l], wd[l]);
            MIPT_HIP_CHECK(hipGetLastError());
        }
        hipLaunchKernelGGL(pack_plain_kernel, dim3(blocks(CH * CH, 256)), dim3(256), 0, stream, w + off_fc, CH * CH, w1b);
        hipLaunchKernelGGL(pack_plain_kernel, dim3(1), dim3(256), 0, stream, w + off_fc + CH * CH + CH, CH * 3, w2b);
        MIPT_HIP_CHECK(hipGetLastError());
        return OK;
    }

    int set_params(const float *host)
    {
        MIPT_HIP_CHECK(hipMemcpyAsync(w, host, nparam * 4, hipMemcpyHostToDevice, stream));
        MIPT_HIP_TRY(pack());
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        return OK;
    }

    // forward over nb samples; rec/idx address the records (idx may be null)
    int forward(const std::int8_t *rec, const int *idx, int nb)
    {
        const int M = nb * HW;
        const int grid = blocks(M, TM);
        hipLaunchKernelGGL(conv_fwd_kernel<true>, dim3(grid), dim3(NT), 0, stream, nullptr, rec, idx, wf[0],
                           w + off_b[0], act[0], M, H, W, kp[0]);
        for (int l = 1; l < 4; ++l)
            hipLaunchKernelGGL(conv_fwd_kernel<false>, dim3(grid), dim3(NT), 0, stream, act[l - 1], nullptr, nullptr,
                               wf[l], w + off_b[l], act[l], M, H, W, kp[l]);
        hipLaunchKernelGGL(pool_fwd_kernel, dim3(nb), dim3(256), 0, stream, act[3], HW, pooled);
        hipLaunchKernelGGL(mlp_fwd_kernel, dim3(nb), dim3(CH), 0, stream, pooled, w1b, w + off_fc + CH * CH, w2b,
                           w + off_fc + CH * CH + CH + CH * 3, h, out);
        MIPT_HIP_CHECK(hipGetLastError());
        return OK;
    }

    // gradients of the batch loss into g (FP32, master layout)
    int backward(const std::int8_t *rec, const int *idx, const float *labels, int nb)
    {
        const int M = nb * HW;
        const int grid = blocks(M, TM);
        hipLaunchKernelGGL(loss_bwd_kernel, dim3(nb), dim3(CH), 0, stream, out, labels, idx, h, w1b, w2b, nb, huber,
                           loss, dout, dh, dpooled);
        hipLaunchKernelGGL(mlp_grad_kernel, dim3(N_FC), dim3(256), 0, stream, pooled, h, dh, dout, nb, g + off_fc);
        hipLaunchKernelGGL(pool_bwd_kernel, dim3(blocks(static_cast<std::size_t>(M) * CH, 256)), dim3(256), 0, stream,
                           dpooled, M, HW, dact);
        MIPT_HIP_CHECK(hipGetLastError());
        for (int l = 3; l >= 0; --l)
        {
            const int krows = 9 * ci[l];
            const int ktiles = blocks(krows, TM);
            // positions per partial: about WG_BLOCKS (first layer) or WG3_BLOCKS blocks in
            // all, at least 8 K steps each
            const int xblocks = l == 0 ? ktiles : 3;
            const int chunk_len = std::max(
                8 * TK, blocks(blocks(M, blocks(l == 0 ? WG_BLOCKS : WG3_BLOCKS, xblocks)), TK) * TK);
            const int nch = blocks(M, chunk_len);
            const int nbp = blocks(M, BROWS);
            hipLaunchKernelGGL(dz_bias_kernel, dim3(nbp), dim3(256), 0, stream, dact, act[l], M, dz, bpart);
            hipLaunchKernelGGL(sum_columns_kernel, dim3(CH), dim3(256), 0, stream, bpart, nbp, CH, g + off_b[l]);
            if (l == 0)
                hipLaunchKernelGGL(conv_bwd_weight_kernel<true>, dim3(ktiles, nch), dim3(NT), 0, stream, nullptr, rec,
                                   idx, dz, partial, M, H, W, krows, chunk_len);
            else
                hipLaunchKernelGGL(conv_bwd_weight3_kernel, dim3(3, nch), dim3(NT), 0, stream, act[l - 1], dz, partial,
                                   M, H, W, chunk_len);
            hipLaunchKernelGGL(sum_partials_kernel, dim3(blocks(krows * CH, 64)), dim3(256), 0, stream, partial, nch,
                               krows * CH, g + off_w[l]);
            if (l > 0)
                hipLaunchKernelGGL(conv_bwd_data_kernel, dim3(grid), dim3(NT), 0, stream, dz, wd[l], dact, M, H, W);
            MIPT_HIP_CHECK(hipGetLastError());
        }
        return OK;
    }

    int adam(float lr)
    {
        ++step;
        const float c1 = 1.f - std::pow(0.9f, static_cast<float>(step));
        const float c2 = 1.f - std::pow(0.999f, static_cast<float>(step));
        hipLaunchKernelGGL(adam_kernel, dim3(blocks(nparam, 256)), dim3(256), 0, stream, w, m, v, g,
                           static_cast<int>(nparam), lr, c1, c2);
        MIPT_HIP_CHECK(hipGetLastError());
        return pack();
    }

    double batch_loss(int nb)
    {
        std::vector<float> l(nb);
        if (hipMemcpyAsync(l.data(), loss, nb * 4, hipMemcpyDeviceToHost, stream) != hipSuccess) return NAN;
        if (hipStreamSynchronize(stream) != hipSuccess) return NAN;
        double s = 0.0;
        for (float x : l) s += x;
        return s / nb;
    }

    int check_indices(const int *host_idx, int n)
    {
        for (int i = 0; i < n; ++i)
            if (host_idx[i] < 0 || host_idx[i] >= data_count)
            {
                set_error("decoder: sample index %d out of the dataset (%d rows)", host_idx[i], data_count);
                return E_ARG;
            }
        return OK;
    }
};

} // namespace mipt_hip

using namespace mipt_hip;

struct mipt_hip_decoder : Decoder
{
};

extern "C" {

int mipt_hip_decoder_create(int height, int width, int batch_max, uint64_t seed, mipt_hip_decoder **out)
{
    if (!out || height < 1 || width < 1 || batch_max < 1 || static_cast<long>(height) * width * batch_max > (1L << 26))
    {
        set_error("decoder_create: bad arguments");
        return E_ARG;
    }
    *out = nullptr;
    auto *d = new (std::nothrow) mipt_hip_decoder();
    if (!d) return E_NOMEM;
    const int st = d->init(height, width, batch_max, seed);
    if (st != OK)
    {
        delete d;
        return st;
    }
    *out = d;
    return OK;
}

int mipt_hip_decoder_destroy(mipt_hip_decoder *d)
{
    delete d;
    return OK;
}

int mipt_hip_decoder_param_count(mipt_hip_decoder *d) { return d ? static_cast<int>(d->nparam) : -1; }

int mipt_hip_decoder_get_params(mipt_hip_decoder *d, float *host)
{
    if (!d || !host) { set_error("decoder_get_params: bad arguments"); return E_ARG; }
    MIPT_HIP_CHECK(hipMemcpyAsync(host, d->w, d->nparam * 4, hipMemcpyDeviceToHost, d->stream));
    MIPT_HIP_CHECK(hipStreamSynchronize(d->stream));
    return OK;
}

int mipt_hip_decoder_set_params(mipt_hip_decoder *d, const float *host)
{
    if (!d || !host) { set_error("decoder_set_params: bad arguments"); return E_ARG; }
    d->step = 0;
    MIPT_HIP_CHECK(hipMemsetAsync(d->m, 0, d->nparam * 4, d->stream));
    MIPT_HIP_CHECK(hipMemsetAsync(d->v, 0, d->nparam * 4, d->stream));
    return d->set_params(host);
}

int mipt_hip_decoder_set_dataset(mipt_hip_decoder *d, const int8_t *host_records, const float *host_labels, int count)
{
    if (!d || !host_records || count < 1) { set_error("decoder_set_dataset: bad arguments"); return E_ARG; }
    if (d->data_rec) (void)hipFree(d->data_rec);
    if (d->data_lab) (void)hipFree(d->data_lab);
    d->data_rec = nullptr;
    d->data_lab = nullptr;
    d->data_count = 0;
    MIPT_HIP_CHECK(hipMalloc(&d->data_rec, static_cast<std::size_t>(count) * d->HW));
    MIPT_HIP_CHECK(hipMalloc(&d->data_lab, static_cast<std::size_t>(count) * 3 * 4));
    MIPT_HIP_CHECK(hipMemcpyAsync(d->data_rec, host_records, static_cast<std::size_t>(count) * d->HW,
                                  hipMemcpyHostToDevice, d->stream));
    if (host_labels)
        MIPT_HIP_CHECK(hipMemcpyAsync(d->data_lab, host_labels, static_cast<std::size_t>(count) * 12,
                                      hipMemcpyHostToDevice, d->stream));
    else
        MIPT_HIP_CHECK(hipMemsetAsync(d->data_lab, 0, static_cast<std::size_t>(count) * 12, d->stream));
    MIPT_HIP_CHECK(hipStreamSynchronize(d->stream));
    d->data_count = count;
    return OK;
}

int mipt_hip_decoder_gradients(mipt_hip_decoder *d, const int *host_idx, int batch, float *host_grad, double *loss)
{
    if (!d || !host_idx || batch < 1 || batch > d->bmax || !host_grad)
    {
        set_error("decoder_gradients: bad arguments");
        return E_ARG;
    }
    MIPT_HIP_TRY(d->check_indices(host_idx, batch));
    MIPT_HIP_CHECK(hipMemcpyAsync(d->d_idx, host_idx, batch * 4, hipMemcpyHostToDevice, d->stream));
    MIPT_HIP_TRY(d->forward(d->data_rec, d->d_idx, batch));
    MIPT_HIP_TRY(d->backward(d->data_rec, d->d_idx, d->data_lab, batch));
    MIPT_HIP_CHECK(hipMemcpyAsync(host_grad, d->g, d->nparam * 4, hipMemcpyDeviceToHost, d->stream));
    const double l = d->batch_loss(batch);
    if (loss) *loss = l;
    return OK;
}

int mipt_hip_decoder_train(mipt_hip_decoder *d, const int *host_idx, int batch, int steps, float lr, double *mean_loss)
{
    if (!d || !host_idx || batch < 1 || batch > d->bmax || steps < 1 || !(lr > 0.f))
    {
        set_error("decoder_train: bad arguments");
        return E_ARG;
    }
    MIPT_HIP_TRY(d->check_indices(host_idx, batch * steps));
    int *dall = nullptr;
    MIPT_HIP_CHECK(hipMalloc(&dall, static_cast<std::size_t>(batch) * steps * 4));
    int st = OK;
    if (hipMemcpyAsync(dall, host_idx, static_cast<std::size_t>(batch) * steps * 4, hipMemcpyHostToDevice, d->stream) !=
        hipSuccess)
        st = E_ARG;
    double acc = 0.0;
    for (int s = 0; s < steps && st == OK; ++s)
    {
        const int *idx = dall + static_cast<std::size_t>(s) * batch;
        if ((st = d->forward(d->data_rec, idx, batch)) != OK) break;
        if ((st = d->backward(d->data_rec, idx, d->data_lab, batch)) != OK) break;
        if ((st = d->adam(lr)) != OK) break;
        if (mean_loss) acc += d->batch_loss(batch);
    }
    (void)hipStreamSynchronize(d->stream);
    (void)hipFree(dall);
    if (mean_loss) *mean_loss = acc / steps;
    return st;
}

int mipt_hip_decoder_predict(mipt_hip_decoder *d, int first, int count, float *host_out)
{
    if (!d || !host_out || first < 0 || count < 1 || first + count > d->data_count)
    {
        set_error("decoder_predict: bad arguments");
        return E_ARG;
    }
    for (int b0 = 0; b0 < count; b0 += d->bmax)
    {
        const int nb = std::min(d->bmax, count - b0);
        MIPT_HIP_TRY(d->forward(d->data_rec + static_cast<std::size_t>(first + b0) * d->HW, nullptr, nb));
        MIPT_HIP_CHECK(hipMemcpyAsync(host_out + static_cast<std::size_t>(b0) * 3, d->out, nb * 12, hipMemcpyDeviceToHost,
                                      d->stream));
    }
    MIPT_HIP_CHECK(hipStreamSynchronize(d->stream));
    return OK;
}

int mipt_hip_decoder_set_loss(mipt_hip_decoder *d, float huber_delta)
{
    if (!d || !(huber_delta >= 0.f)) { set_error("decoder_set_loss: bad arguments"); return E_ARG; }
    d->huber = huber_delta;
    return OK;
}

int mipt_hip_decoder_synchronize(mipt_hip_decoder *d)
{
    if (!d) { set_error("decoder_synchronize: null decoder"); return E_ARG; }
    MIPT_HIP_CHECK(hipStreamSynchronize(d->stream));
    return OK;
}

} // extern "C"
