// Reduced density matrices, entanglement entropies, TMI and the probe rho1.
//
// Reference behaviour (paths relative to the reference repository root):
//   cpp/include/mipt/tmi/linalg.hpp:885-1026     kept/env offsets, Schmidt gather
//   cpp/include/mipt/tmi/linalg.hpp:125-141, 201-208  entropy cut-offs, Renyi-2
//   cpp/include/mipt/cuda/schmidt_gram.cuh:270-315, 538-724  gather + herk
//   cpp/src/cuda/tmi_cuda_svd.cu:663-798, 801-851 entropy C ABI
//   cpp/src/cuda/tmi_cuda_rdm.cu:59-144          RDM export (row-major interleaved)
//   cpp/include/mipt/tmi/compute.hpp:89-113, 186-227  TMI regions and formula
//   cpp/src/cuda/ancilla_rdm.cu:228-297          probe rho1
//
// gfx950 design: the Schmidt matrix M (2^kept x 2^env) is gathered once into
// split real/imaginary FP32 planes; rho = M M^H is computed as the two real
// GEMMs rho_r = Ar Ar^T + Ai Ai^T and rho_i = Ai Ar^T - Ar Ai^T on MFMA,
// lower-triangle 64x64 tiles only, split-K, then a fixed-order FP64 split-K
// reduction.  The products run on FP64 MFMA (exact products of the complex64
// inputs, FP64 accumulation) by default, or on FP32 MFMA with FP64 promotion
// every 16-deep K step.  Eigenvalues of the FP64 rho come from the library's
// batched Householder + bisection kernels (eigen.hip) for batches and for
// single RDMs up to 512 x 512, from rocSOLVER zheevd for larger single RDMs
// (or everywhere when selected); the -sum lambda log2 lambda is FP64.
#include "batch.hpp"
#include "state_ops.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

namespace mipt_hip
{
namespace
{
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int GT = 64;         // output tile
constexpr int GK = 16;         // K step staged through LDS
constexpr int KCHUNK = 16;     // FP32 accumulation depth before FP64 promotion

// M[r, c] = psi[env_off(c) | row_off(r)], written as planes Ar[r][c], Ai[r][c].
template <typename C>
__global__ void gather_kernel(const C *__restrict__ psi, const std::uint64_t *__restrict__ row_off, int kept, int env,
                              std::uint64_t env_mask, float *__restrict__ ar, float *__restrict__ ai)
{
    const std::uint64_t cols = std::uint64_t{1} << env;
    const std::uint64_t total = cols << kept;
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
    for (std::uint64_t id = blockIdx.x * static_cast<std::uint64_t>(blockDim.x) + threadIdx.x; id < total; id += stride)
    {
        const std::uint64_t r = id >> env;
        const std::uint64_t c = id & (cols - 1);
        const C v = psi[pdep64(c, env_mask) | row_off[r]];
        ar[id] = static_cast<float>(v.x);
        ai[id] = static_cast<float>(v.y);
    }
}

// Four consecutive K entries k0..k0+3 of row g of a plane (row length K)
// into dst, entries past k1 or rows past m as 0: one 16-byte load when all
// four are in range (K a power of two of at least 4 keeps it aligned).
__device__ __forceinline__ void stage_k4(float *dst, const float *__restrict__ plane, int g, int m, int K, int k0, int k1)
{
    if (g < m && k0 + 3 < k1)
    {
        const float4 v = *reinterpret_cast<const float4 *>(plane + static_cast<std::size_t>(g) * K + k0);
        dst[0] = v.x;
        dst[1] = v.y;
        dst[2] = v.z;
        dst[3] = v.w;
    }
    else
        for (int e = 0; e < 4; ++e)
            dst[e] = (g < m && k0 + e < k1) ? plane[static_cast<std::size_t>(g) * K + k0 + e] : 0.f;
}

// Lower-triangle tile (bi >= bj) of rho over K range [k0, k1) of split `sp`;
// FP64 partial written to part[sp][tile][64][64] as (re, im).
__global__ void __launch_bounds__(256) gram_mfma_kernel(const float *__restrict__ ar, const float *__restrict__ ai,
                                                        int m, int K, int ksplit, int ntiles_tri,
                                                        double2 *__restrict__ part)
{
    __shared__ float sA_r[GT][GK + 1], sA_i[GT][GK + 1], sB_r[GT][GK + 1], sB_i[GT][GK + 1];
    const int tile = blockIdx.x % ntiles_tri;
    const int sp = blockIdx.x / ntiles_tri;
    // tile -> (bi, bj) with bj <= bi
    int bi = static_cast<int>((sqrt(8.0 * tile + 1.0) - 1.0) / 2.0);
    while ((bi + 1) * (bi + 2) / 2 <= tile) ++bi;
    while (bi * (bi + 1) / 2 > tile) --bi;
    const int bj = tile - bi * (bi + 1) / 2;
    const int kper = (K + ksplit - 1) / ksplit;
    const int k0 = sp * kper, k1 = min(K, k0 + kper);

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wi = wave >> 1, wj = wave & 1;  // 32x32 quadrant of the 64x64 tile
    floatx4 accr[2][2], acci[2][2];
    double dr[2][2][4], di[2][2][4];
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y)
        {
            accr[x][y] = floatx4{0, 0, 0, 0};
            acci[x][y] = floatx4{0, 0, 0, 0};
            for (int e = 0; e < 4; ++e) dr[x][y][e] = di[x][y][e] = 0.0;
        }
    const int lrow = threadIdx.x >> 2;       // 0..63
    const int lk = (threadIdx.x & 3) * 4;    // 0,4,8,12
    int since = 0;
    for (int kk = k0; kk < k1; kk += GK)
    {
        // stage 64 rows x 16 k of the i-block and j-block
        const int gi = bi * GT + lrow, gj = bj * GT + lrow;
        stage_k4(&sA_r[lrow][lk], ar, gi, m, K, kk + lk, k1);
        stage_k4(&sA_i[lrow][lk], ai, gi, m, K, kk + lk, k1);
        stage_k4(&sB_r[lrow][lk], ar, gj, m, K, kk + lk, k1);
        stage_k4(&sB_i[lrow][lk], ai, gj, m, K, kk + lk, k1);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK; ks += 4)
        {
            const int kq = ks + (lane >> 4);
#pragma unroll
            for (int x = 0; x < 2; ++x)
            {
                const int ra = wi * 32 + x * 16 + (lane & 15);
                const float a_r = sA_r[ra][kq], a_i = sA_i[ra][kq];
#pragma unroll
                for (int y = 0; y < 2; ++y)
                {
                    const int rb = wj * 32 + y * 16 + (lane & 15);
                    const float b_r = sB_r[rb][kq], b_i = sB_i[rb][kq];
                    accr[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_r, b_r, accr[x][y], 0, 0, 0);
                    accr[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_i, b_i, accr[x][y], 0, 0, 0);
                    acci[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_i, b_r, acci[x][y], 0, 0, 0);
                    acci[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(-a_r, b_i, acci[x][y], 0, 0, 0);
                }
            }
        }
        __syncthreads();
        since += GK;
        if (since >= KCHUNK || kk + GK >= k1)
        {
            for (int x = 0; x < 2; ++x)
                for (int y = 0; y < 2; ++y)
                    for (int e = 0; e < 4; ++e)
                    {
                        dr[x][y][e] += static_cast<double>(accr[x][y][e]);
                        di[x][y][e] += static_cast<double>(acci[x][y][e]);
                        accr[x][y][e] = 0.f;
                        acci[x][y][e] = 0.f;
                    }
            since = 0;
        }
    }
    // D layout of 16x16x4: lane holds rows 4*(lane/16)+e, column lane%16
    double2 *out = part + (static_cast<std::size_t>(sp) * ntiles_tri + tile) * GT * GT;
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y)
            for (int e = 0; e < 4; ++e)
            {
                const int r = wi * 32 + x * 16 + 4 * (lane >> 4) + e;
                const int c = wj * 32 + y * 16 + (lane & 15);
                out[r * GT + c] = make_double2(dr[x][y][e], di[x][y][e]);
            }
}

// Same tiling with exact products: the complex64 inputs are widened to FP64
// and multiplied on FP64 MFMA (v_mfma_f64_16x16x4f64) with FP64 accumulation,
// so rho carries only the rounding of the complex64 state itself.
__global__ void __launch_bounds__(256) gram_mfma64_kernel(const float *__restrict__ ar, const float *__restrict__ ai,
                                                          int m, int K, int ksplit, int ntiles_tri,
                                                          double2 *__restrict__ part)
{
    __shared__ float sA_r[GT][GK + 1], sA_i[GT][GK + 1], sB_r[GT][GK + 1], sB_i[GT][GK + 1];
    const int tile = blockIdx.x % ntiles_tri;
    const int sp = blockIdx.x / ntiles_tri;
    int bi = static_cast<int>((sqrt(8.0 * tile + 1.0) - 1.0) / 2.0);
    while ((bi + 1) * (bi + 2) / 2 <= tile) ++bi;
    while (bi * (bi + 1) / 2 > tile) --bi;
    const int bj = tile - bi * (bi + 1) / 2;
    const int kper = (K + ksplit - 1) / ksplit;
    const int k0 = sp * kper, k1 = min(K, k0 + kper);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wi = wave >> 1, wj = wave & 1;
    doublex4 accr[2][2], acci[2][2];
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y)
        {
            accr[x][y] = doublex4{0, 0, 0, 0};
            acci[x][y] = doublex4{0, 0, 0, 0};
        }
    const int lrow = threadIdx.x >> 2;
    const int lk = (threadIdx.x & 3) * 4;
    for (int kk = k0; kk < k1; kk += GK)
    {
        const int gi = bi * GT + lrow, gj = bj * GT + lrow;
        stage_k4(&sA_r[lrow][lk], ar, gi, m, K, kk + lk, k1);
        stage_k4(&sA_i[lrow][lk], ai, gi, m, K, kk + lk, k1);
        stage_k4(&sB_r[lrow][lk], ar, gj, m, K, kk + lk, k1);
        stage_k4(&sB_i[lrow][lk], ai, gj, m, K, kk + lk, k1);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK; ks += 4)
        {
            const int kq = ks + (lane >> 4);
#pragma unroll
            for (int x = 0; x < 2; ++x)
            {
                const int ra = wi * 32 + x * 16 + (lane & 15);
                const double a_r = sA_r[ra][kq], a_i = sA_i[ra][kq];
#pragma unroll
                for (int y = 0; y < 2; ++y)
                {
                    const int rb = wj * 32 + y * 16 + (lane & 15);
                    const double b_r = sB_r[rb][kq], b_i = sB_i[rb][kq];
                    accr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r, b_r, accr[x][y], 0, 0, 0);
                    accr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i, b_i, accr[x][y], 0, 0, 0);
                    acci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i, b_r, acci[x][y], 0, 0, 0);
                    acci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a_r, b_i, acci[x][y], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D layout of f64 16x16x4: lane holds rows lane/16 + 4*e, column lane%16
    double2 *out = part + (static_cast<std::size_t>(sp) * ntiles_tri + tile) * GT * GT;
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y)
            for (int e = 0; e < 4; ++e)
            {
                const int r = wi * 32 + x * 16 + (lane >> 4) + 4 * e;
                const int c = wj * 32 + y * 16 + (lane & 15);
                out[r * GT + c] = make_double2(accr[x][y][e], acci[x][y][e]);
            }
}

// Sum the split-K partials in order into rho (m x m, row-major double2),
// mirroring the lower triangle and forcing a real diagonal.
__global__ void gram_reduce_kernel(const double2 *__restrict__ part, int m, int ksplit, int ntiles_tri,
                                   double2 *__restrict__ rho)
{
    const std::size_t id = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
    if (id >= static_cast<std::size_t>(m) * m) return;
    const int r = static_cast<int>(id / m), c = static_cast<int>(id % m);
    const bool lower = r >= c;
    const int rr = lower ? r : c, cc = lower ? c : r;
    const int bi = rr / GT, bj = cc / GT;
    const int tile = bi * (bi + 1) / 2 + bj;
    const int lr = rr % GT, lc = cc % GT;
    double sr = 0.0, si = 0.0;
    for (int sp = 0; sp < ksplit; ++sp)
    {
        const double2 v = part[(static_cast<std::size_t>(sp) * ntiles_tri + tile) * GT * GT + lr * GT + lc];
        sr += v.x;
        si += v.y;
    }
    if (r == c) si = 0.0;
    rho[id] = lower ? make_double2(sr, si) : make_double2(sr, -si);
}

// Batched Schmidt gather for many (state, kept set) items of one shape:
// item t's planes at t * m * K, M_t[r, c] = psi_t[pdep(c, env_mask_t) | row_off_t[r]].
__global__ void gather_batch_kernel(const float2 *const *__restrict__ psis, const std::uint64_t *__restrict__ row_off,
                                    const std::uint64_t *__restrict__ env_mask, int kept, int env, int count,
                                    float *__restrict__ ar, float *__restrict__ ai)
{
    const std::uint64_t per = std::uint64_t{1} << (kept + env);
    const std::uint64_t total = per * static_cast<std::uint64_t>(count);
    const std::uint64_t cols = std::uint64_t{1} << env;
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
    for (std::uint64_t id = blockIdx.x * static_cast<std::uint64_t>(blockDim.x) + threadIdx.x; id < total; id += stride)
    {
        const std::uint64_t t = id >> (kept + env);
        const std::uint64_t loc = id & (per - 1);
        const std::uint64_t r = loc >> env, c = loc & (cols - 1);
        const float2 v = psis[t][pdep64(c, env_mask[t]) | row_off[(t << kept) + r]];
        ar[id] = v.x;
        ai[id] = v.y;
    }
}

// Batched Gram on FP64 MFMA for the library's eigensolver: item blockIdx.y,
// lower tile (bi >= bj) of rho = M M^H over the full K (no split), written as
// the UPPER triangle of the row-major rho, rho[a][b] for a <= b, i.e. operand
// A from the rows of tile bj and B from the rows of tile bi (each wave store
// covers 16 consecutive columns).  The eigensolver reads that triangle as
// the lower triangle of the column-major conj(rho), which has the same
// spectrum.  Exact products of the complex64 inputs, FP64 sums, real diagonal.
template <int TS>
__global__ void __launch_bounds__(256) gram_batch64_kernel(const float *__restrict__ ar_all,
                                                           const float *__restrict__ ai_all, int m, int K,
                                                           int ntiles_tri, double2 *__restrict__ rho_all)
{
    // TS x TS output tiles (64, or 32 for RDMs of at most 32 rows); each wave
    // takes a quadrant of QS x QS in 16 x 16 MFMA blocks
    constexpr int QS = TS / 2, NX = QS / 16;
    __shared__ float sA_r[TS][GK + 1], sA_i[TS][GK + 1], sB_r[TS][GK + 1], sB_i[TS][GK + 1];
    const int tile = blockIdx.x;
    const std::size_t item = blockIdx.y;
    const float *ar = ar_all + item * static_cast<std::size_t>(m) * K;
    const float *ai = ai_all + item * static_cast<std::size_t>(m) * K;
    double2 *rho = rho_all + item * static_cast<std::size_t>(m) * m;
    int bi = static_cast<int>((sqrt(8.0 * tile + 1.0) - 1.0) / 2.0);
    while ((bi + 1) * (bi + 2) / 2 <= tile) ++bi;
    while (bi * (bi + 1) / 2 > tile) --bi;
    const int bj = tile - bi * (bi + 1) / 2;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wi = wave >> 1, wj = wave & 1;
    doublex4 accr[NX][NX], acci[NX][NX];
    for (int x = 0; x < NX; ++x)
        for (int y = 0; y < NX; ++y)
        {
            accr[x][y] = doublex4{0, 0, 0, 0};
            acci[x][y] = doublex4{0, 0, 0, 0};
        }
    const int lrow = threadIdx.x >> 2;
    const int lk = (threadIdx.x & 3) * 4;
    const int ga = bj * TS + lrow, gb = bi * TS + lrow;
    for (int kk = 0; kk < K; kk += GK)
    {
        if (lrow < TS)
        {
            stage_k4(&sA_r[lrow][lk], ar, ga, m, K, kk + lk, K);
            stage_k4(&sA_i[lrow][lk], ai, ga, m, K, kk + lk, K);
            stage_k4(&sB_r[lrow][lk], ar, gb, m, K, kk + lk, K);
            stage_k4(&sB_i[lrow][lk], ai, gb, m, K, kk + lk, K);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK; ks += 4)
        {
            const int kq = ks + (lane >> 4);
#pragma unroll
            for (int x = 0; x < NX; ++x)
            {
                const int ra = wi * QS + x * 16 + (lane & 15);
                const double a_r = sA_r[ra][kq], a_i = sA_i[ra][kq];
#pragma unroll
                for (int y = 0; y < NX; ++y)
                {
                    const int rb = wj * QS + y * 16 + (lane & 15);
                    const double b_r = sB_r[rb][kq], b_i = sB_i[rb][kq];
                    accr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r, b_r, accr[x][y], 0, 0, 0);
                    accr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i, b_i, accr[x][y], 0, 0, 0);
                    acci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i, b_r, acci[x][y], 0, 0, 0);
                    acci[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a_r, b_i, acci[x][y], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    for (int x = 0; x < NX; ++x)
        for (int y = 0; y < NX; ++y)
            for (int e = 0; e < 4; ++e)
            {
                const int a = bj * TS + wi * QS + x * 16 + (lane >> 4) + 4 * e;
                const int b = bi * TS + wj * QS + y * 16 + (lane & 15);
                if (a < m && b < m && a <= b)
                    rho[static_cast<std::size_t>(a) * m + b] = make_double2(accr[x][y][e], a == b ? 0.0 : acci[x][y][e]);
            }
}

// FP64 Gram for complex128 states (drop-in _f64 path): rho[r][c] = sum_k M[r,k] conj(M[c,k]).
__global__ void gram_f64_kernel(const double2 *__restrict__ psi, const std::uint64_t *__restrict__ row_off, int m,
                                int env, std::uint64_t env_mask, double2 *__restrict__ rho)
{
    const std::size_t id = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
    if (id >= static_cast<std::size_t>(m) * m) return;
    const int r = static_cast<int>(id / m), c = static_cast<int>(id % m);
    if (c > r) return;
    const std::uint64_t cols = std::uint64_t{1} << env;
    double sr = 0.0, si = 0.0;
    for (std::uint64_t k = 0; k < cols; ++k)
    {
        const std::uint64_t e = pdep64(k, env_mask);
        const double2 a = psi[e | row_off[r]], b = psi[e | row_off[c]];
        sr += a.x * b.x + a.y * b.y;
        si += a.y * b.x - a.x * b.y;
    }
    if (r == c) si = 0.0;
    rho[static_cast<std::size_t>(r) * m + c] = make_double2(sr, si);
    rho[static_cast<std::size_t>(c) * m + r] = make_double2(sr, -si);
}

__global__ void purity_partial_kernel(const double2 *__restrict__ rho, std::size_t n2, double *__restrict__ part)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n2;
         i += static_cast<std::size_t>(gridDim.x) * blockDim.x)
    {
        const double2 v = rho[i];
        acc += v.x * v.x + v.y * v.y;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

template <typename C>
__global__ void rho1_kernel(const C *__restrict__ psi, int n, int q, double *__restrict__ part)
{
    __shared__ double red[4][256];
    const std::uint64_t env = std::uint64_t{1} << (n - 1);
    const std::uint64_t low = (std::uint64_t{1} << q) - 1;
    double a00 = 0, a11 = 0, re = 0, im = 0;
    for (std::uint64_t e = blockIdx.x * static_cast<std::uint64_t>(blockDim.x) + threadIdx.x; e < env;
         e += static_cast<std::uint64_t>(gridDim.x) * blockDim.x)
    {
        const std::uint64_t i0 = (e & low) | ((e & ~low) << 1);
        const C a = psi[i0], b = psi[i0 | (std::uint64_t{1} << q)];
        const double ar = a.x, ai = a.y, br = b.x, bi = b.y;
        a00 += ar * ar + ai * ai;
        a11 += br * br + bi * bi;
        re += ar * br + ai * bi;
        im += ai * br - ar * bi;
    }
    red[0][threadIdx.x] = a00;
    red[1][threadIdx.x] = a11;
    red[2][threadIdx.x] = re;
    red[3][threadIdx.x] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
            for (int v = 0; v < 4; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int v = 0; v < 4; ++v) part[blockIdx.x * 4 + v] = red[v][0];
}

// ------------------------------------------------------------ host helpers
} // namespace

struct Workspace
{
    float *ar = nullptr, *ai = nullptr;
    std::size_t plane_cap = 0, plane_cap_i = 0;
    double2 *part = nullptr;
    std::size_t part_cap = 0;
    double2 *rho = nullptr;
    std::size_t rho_cap = 0;
    double *evals = nullptr, *e = nullptr;
    std::size_t ev_cap = 0, e_cap = 0;
    int *info = nullptr;
    std::uint64_t *row_off = nullptr;
    std::size_t row_cap = 0;
    double *red = nullptr;
    rocblas_handle handle = nullptr;
    // batched eigensolves: matrices, eigenvalues, off-diagonals, infos
    double2 *rho_b = nullptr;
    std::size_t rho_b_cap = 0;
    double *ev_b = nullptr, *e_b = nullptr;
    std::size_t ev_b_cap = 0, e_b_cap = 0;
    int *info_b = nullptr;
    std::size_t info_b_cap = 0;
    // the project's batched eigensolver: panel workspace, tridiagonals, entropies
    double2 *eig_w = nullptr;
    std::size_t eig_w_cap = 0;
    double *eig_d = nullptr, *eig_e = nullptr, *eig_s = nullptr;
    std::size_t eig_d_cap = 0, eig_e_cap = 0, eig_s_cap = 0;
    // batched gather tables: row offsets, environment masks, state pointers
    std::uint64_t *rowb = nullptr, *emb = nullptr, *psib = nullptr;
    std::size_t rowb_cap = 0, emb_cap = 0, psib_cap = 0;
};

Workspace *workspace_create() { return new Workspace(); }

// Frees every buffer and the rocSOLVER handle (the caller has synchronised
// the streams that used them).
void workspace_destroy(Workspace *w)
{
    if (!w) return;
    void *bufs[] = {w->ar, w->ai, w->part, w->rho, w->evals, w->e, w->info, w->row_off, w->red, w->rho_b, w->ev_b,
                    w->e_b, w->info_b, w->eig_w, w->eig_d, w->eig_e, w->eig_s, w->rowb, w->emb, w->psib};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (w->handle) (void)rocblas_destroy_handle(w->handle);
    delete w;
}

namespace
{

// Workspace of the context-free drop-in entry points (the reference's
// signatures carry no context): one per calling thread for the process's
// lifetime (never freed: the reference keeps its workspaces thread_local too,
// and freeing at thread exit could run after the HIP runtime is gone).  A
// batch passes its own workspace instead.
thread_local Workspace *t_ws = nullptr;

Workspace &ws()
{
    if (!t_ws) t_ws = new Workspace();
    return *t_ws;
}

template <typename T>
int grow(T *&p, std::size_t &cap, std::size_t need)
{
    if (need <= cap) return OK;
    if (p) MIPT_HIP_CHECK(hipFree(p));
    p = nullptr;
    MIPT_HIP_CHECK(hipMalloc(&p, need * sizeof(T)));
    cap = need;
    return OK;
}

int validate_kept(int n, const int *kept, int count, std::uint64_t &kmask)
{
    kmask = 0;
    if (!kept || count < 0 || count > 30 || n < 1 || n > 34 || count > n) return E_ARG;
    for (int j = 0; j < count; ++j)
    {
        if (kept[j] < 0 || kept[j] >= n || ((kmask >> kept[j]) & 1u)) return E_ARG;
        kmask |= std::uint64_t{1} << kept[j];
    }
    return OK;
}

// Gram arithmetic of the complex64 path: 64 (default) = FP64 MFMA with exact
// products; 32 = FP32 MFMA promoted to FP64 every 16-deep K step.  Chosen by
// mipt_hip_set_gram_precision, else MIPT_HIP_GRAM_PRECISION (32 or 64; any
// other value is refused).  The FP32 form measured up to 1e-5 bits of entropy
// error on half-chain RDMs, so it is not the default.
std::atomic<int> g_gram_bits{0};  // 0: not chosen yet

int gram_precision(int &bits)
{
    int b = g_gram_bits.load();
    if (b == 0)
    {
        const char *e = std::getenv("MIPT_HIP_GRAM_PRECISION");
        if (!e || !*e || std::strcmp(e, "64") == 0) b = 64;
        else if (std::strcmp(e, "32") == 0) b = 32;
        else
        {
            set_error("MIPT_HIP_GRAM_PRECISION must be 32 or 64, got '%s'", e);
            return E_ARG;
        }
        int expect = 0;
        g_gram_bits.compare_exchange_strong(expect, b);
        b = g_gram_bits.load();
    }
    bits = b;
    return OK;
}

// Eigenvalue path of the entropies: 0 (default) = the library's batched
// Householder + bisection kernels (eigen.hip) for RDMs up to EIG_MAX_N;
// 1 = rocSOLVER zheevd for every size.  Chosen by mipt_hip_set_eigensolver.
std::atomic<int> g_eig_solver{0};

// Entropies (bits) of `count` RDMs already in w.rho_b (stride m*m) with the
// library's batched eigensolver; results to host out[count].
int own_entropies(Workspace &w, int m, int count, hipStream_t stream, double2 *mats, double *out)
{
    MIPT_HIP_TRY(grow(w.eig_w, w.eig_w_cap, static_cast<std::size_t>(count) * m * 32));
    MIPT_HIP_TRY(grow(w.eig_d, w.eig_d_cap, static_cast<std::size_t>(count) * m));
    MIPT_HIP_TRY(grow(w.eig_e, w.eig_e_cap, static_cast<std::size_t>(count) * m));
    MIPT_HIP_TRY(grow(w.eig_s, w.eig_s_cap, static_cast<std::size_t>(count)));
    MIPT_HIP_TRY(launch_hermitian_entropies(mats, m, static_cast<std::size_t>(m) * m, count, w.eig_w, w.eig_d, w.eig_e,
                                            w.eig_s, nullptr, stream));
    MIPT_HIP_CHECK(hipMemcpyAsync(out, w.eig_s, static_cast<std::size_t>(count) * sizeof(double), hipMemcpyDeviceToHost,
                                  stream));
    MIPT_HIP_CHECK(hipStreamSynchronize(stream));
    return OK;
}

// Optional timing of the Gram kernel (HIP events on the calling stream): its
// average duration and algorithmic FLOPs, 4 m (m + 1) K for the lower
// triangle of M M^H (complex, as real products), for the MFMA roofline.
struct GramTiming
{
    std::mutex mu;
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<double> flops;
};
GramTiming &gram_timing()
{
    static GramTiming t;
    return t;
}

int gram_timing_begin(hipStream_t stream, double flops, int &slot)
{
    GramTiming &g = gram_timing();
    std::lock_guard<std::mutex> lock(g.mu);
    slot = -1;
    if (!g.on) return OK;
    hipEvent_t a, b;
    MIPT_HIP_CHECK(hipEventCreate(&a));
    MIPT_HIP_CHECK(hipEventCreate(&b));
    MIPT_HIP_CHECK(hipEventRecord(a, stream));
    g.ev.emplace_back(a, b);
    g.flops.push_back(flops);
    slot = static_cast<int>(g.ev.size()) - 1;
    return OK;
}

int gram_timing_end(hipStream_t stream, int slot)
{
    if (slot < 0) return OK;
    GramTiming &g = gram_timing();
    std::lock_guard<std::mutex> lock(g.mu);
    MIPT_HIP_CHECK(hipEventRecord(g.ev[slot].second, stream));
    return OK;
}

// rho (FP64, device, row-major m x m) of the kept modes of psi.
int build_rho(Workspace &w, const void *psi, bool f64, int n, const int *kept, int count, hipStream_t stream, int &m_out,
              double2 *dst = nullptr)
{
    std::uint64_t kmask;
    if (validate_kept(n, kept, count, kmask) != OK)
    {
        set_error("invalid kept-mode list");
        return E_ARG;
    }
    const int m = 1 << count;
    const int env = n - count;
    const std::uint64_t env_mask = ((n >= 64) ? ~std::uint64_t{0} : ((std::uint64_t{1} << n) - 1)) & ~kmask;
    std::vector<std::uint64_t> roff(m, 0);
    for (int r = 0; r < m; ++r)
        for (int j = 0; j < count; ++j)
            if ((r >> j) & 1) roff[r] |= std::uint64_t{1} << kept[j];
    MIPT_HIP_TRY(grow(w.row_off, w.row_cap, static_cast<std::size_t>(m)));
    MIPT_HIP_CHECK(hipMemcpyAsync(w.row_off, roff.data(), m * 8, hipMemcpyHostToDevice, stream));
    if (!dst)
    {
        MIPT_HIP_TRY(grow(w.rho, w.rho_cap, static_cast<std::size_t>(m) * m));
        dst = w.rho;
    }
    if (f64)
    {
        const std::size_t mm = static_cast<std::size_t>(m) * m;
        hipLaunchKernelGGL(gram_f64_kernel, dim3(static_cast<unsigned>((mm + 255) / 256)), dim3(256), 0, stream,
                           static_cast<const double2 *>(psi), w.row_off, m, env, env_mask, dst);
        MIPT_HIP_CHECK(hipGetLastError());
    }
    else
    {
        const std::size_t K = std::size_t{1} << env;
        const std::size_t plane = static_cast<std::size_t>(m) * K;
        MIPT_HIP_TRY(grow(w.ar, w.plane_cap, plane));
        MIPT_HIP_TRY(grow(w.ai, w.plane_cap_i, plane));
        const unsigned gblocks = static_cast<unsigned>(std::min<std::size_t>((plane + 255) / 256, 65536));
        hipLaunchKernelGGL(gather_kernel<float2>, dim3(gblocks), dim3(256), 0, stream, static_cast<const float2 *>(psi),
                           w.row_off, count, env, env_mask, w.ar, w.ai);
        MIPT_HIP_CHECK(hipGetLastError());
        const int nb = (m + GT - 1) / GT;
        const int ntri = nb * (nb + 1) / 2;
        int ksplit = 1;
        while (ntri * ksplit < 1024 && static_cast<std::size_t>(ksplit) * 2 * KCHUNK <= K) ksplit *= 2;
        MIPT_HIP_TRY(grow(w.part, w.part_cap, static_cast<std::size_t>(ksplit) * ntri * GT * GT));
        int gbits = 64;
        MIPT_HIP_TRY(gram_precision(gbits));
        int slot = -1;
        MIPT_HIP_TRY(gram_timing_begin(stream, 4.0 * m * (m + 1.0) * static_cast<double>(K), slot));
        if (gbits == 32)
            hipLaunchKernelGGL(gram_mfma_kernel, dim3(static_cast<unsigned>(ntri * ksplit)), dim3(256), 0, stream,
                               w.ar, w.ai, m, static_cast<int>(K), ksplit, ntri, w.part);
        else
            hipLaunchKernelGGL(gram_mfma64_kernel, dim3(static_cast<unsigned>(ntri * ksplit)), dim3(256), 0, stream,
                               w.ar, w.ai, m, static_cast<int>(K), ksplit, ntri, w.part);
        MIPT_HIP_CHECK(hipGetLastError());
        MIPT_HIP_TRY(gram_timing_end(stream, slot));
        const std::size_t mm = static_cast<std::size_t>(m) * m;
        hipLaunchKernelGGL(gram_reduce_kernel, dim3(static_cast<unsigned>((mm + 255) / 256)), dim3(256), 0, stream,
                           w.part, m, ksplit, ntri, dst);
        MIPT_HIP_CHECK(hipGetLastError());
    }
    m_out = m;
    return OK;
}

// Eigenvalues of w.rho (m x m) into host vector.
int eigvalsh(Workspace &w, int m, hipStream_t stream, std::vector<double> &lam)
{
    if (!w.handle)
    {
        const rocblas_status s = rocblas_create_handle(&w.handle);
        if (s != rocblas_status_success)
        {
            set_error("rocblas_create_handle failed (%d)", static_cast<int>(s));
            return E_SOLVER + static_cast<int>(s);
        }
    }
    rocblas_set_stream(w.handle, stream);
    MIPT_HIP_TRY(grow(w.evals, w.ev_cap, static_cast<std::size_t>(m)));
    MIPT_HIP_TRY(grow(w.e, w.e_cap, static_cast<std::size_t>(m)));
    if (!w.info) MIPT_HIP_CHECK(hipMalloc(&w.info, sizeof(int)));
    const rocblas_status s = rocsolver_zheevd(w.handle, rocblas_evect_none, rocblas_fill_lower, m,
                                              reinterpret_cast<rocblas_double_complex *>(w.rho), m, w.evals, w.e, w.info);
    if (s != rocblas_status_success)
    {
        set_error("rocsolver_zheevd failed (%d)", static_cast<int>(s));
        return E_SOLVER + static_cast<int>(s);
    }
    int info = 0;
    lam.resize(m);
    MIPT_HIP_CHECK(hipMemcpyAsync(lam.data(), w.evals, m * 8, hipMemcpyDeviceToHost, stream));
    MIPT_HIP_CHECK(hipMemcpyAsync(&info, w.info, 4, hipMemcpyDeviceToHost, stream));
    MIPT_HIP_CHECK(hipStreamSynchronize(stream));
    if (info != 0)
    {
        set_error("FP64 zheevd did not converge (info=%d)", info);
        return E_EIG + info;
    }
    return OK;
}

// cpp/include/mipt/tmi/linalg.hpp:125-141
double entropy_from_eigs(const std::vector<double> &lam)
{
    double s = 0.0;
    for (double l : lam)
        if (l > 1.0e-15) s += -l * std::log2(l);
    return std::abs(s) < 1.0e-12 ? 0.0 : s;
}

int entropy_impl(Workspace &w, const void *psi, bool f64, int n, const int *kept, int count, hipStream_t stream, int order,
                 double *out)
{
    if (!psi || !out)
    {
        set_error("entropy: null argument");
        return E_ARG;
    }
    if (count == 0 || count == n)
    {
        std::uint64_t km;
        if (validate_kept(n, kept, count, km) != OK && count != 0)
        {
            set_error("invalid kept-mode list");
            return E_ARG;
        }
        *out = 0.0;
        return OK;
    }
    int m = 0;
    MIPT_HIP_TRY(build_rho(w, psi, f64, n, kept, count, stream, m));
    if (order == 2)
    {
        const std::size_t n2 = static_cast<std::size_t>(m) * m;
        const int blocks = static_cast<int>(std::min<std::size_t>(1024, (n2 + 255) / 256));
        if (!w.red) MIPT_HIP_CHECK(hipMalloc(&w.red, 1024 * sizeof(double)));
        hipLaunchKernelGGL(purity_partial_kernel, dim3(blocks), dim3(256), 0, stream, w.rho, n2, w.red);
        MIPT_HIP_CHECK(hipGetLastError());
        std::vector<double> part(blocks);
        MIPT_HIP_CHECK(hipMemcpyAsync(part.data(), w.red, blocks * 8, hipMemcpyDeviceToHost, stream));
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        double purity = 0.0;
        for (double v : part) purity += v;  // fixed block order
        // cpp/include/mipt/tmi/linalg.hpp:201-208
        if (!(purity > 0.0) || purity >= 1.0) *out = 0.0;
        else
        {
            const double s = -std::log2(purity);
            *out = std::abs(s) < 1.0e-12 ? 0.0 : s;
        }
        return OK;
    }
    if (g_eig_solver.load() == 0 && m <= 512)
    {
        // one workgroup per matrix: cheaper than rocSOLVER's launch sequence up to this size
        return own_entropies(w, m, 1, stream, w.rho, out);
    }
    std::vector<double> lam;
    MIPT_HIP_TRY(eigvalsh(w, m, stream, lam));
    *out = entropy_from_eigs(lam);
    return OK;
}

template <typename C>
int rho1_impl(Workspace &w, const void *psi, int n, int q, hipStream_t stream, double *out4)
{
    if (!psi || !out4 || n < 1 || n >= 63 || q < 0 || q >= n)
    {
        set_error("rho1: bad arguments");
        return E_ARG;
    }
    if (!w.red) MIPT_HIP_CHECK(hipMalloc(&w.red, 1024 * sizeof(double)));
    const std::uint64_t env = std::uint64_t{1} << (n - 1);
    const int blocks = static_cast<int>(std::min<std::uint64_t>(256, (env + 255) / 256));
    hipLaunchKernelGGL(rho1_kernel<C>, dim3(blocks), dim3(256), 0, stream, static_cast<const C *>(psi), n, q, w.red);
    MIPT_HIP_CHECK(hipGetLastError());
    std::vector<double> part(static_cast<std::size_t>(blocks) * 4);
    MIPT_HIP_CHECK(hipMemcpyAsync(part.data(), w.red, part.size() * 8, hipMemcpyDeviceToHost, stream));
    MIPT_HIP_CHECK(hipStreamSynchronize(stream));
    for (int v = 0; v < 4; ++v) out4[v] = 0.0;
    for (int b = 0; b < blocks; ++b)
        for (int v = 0; v < 4; ++v) out4[v] += part[static_cast<std::size_t>(b) * 4 + v];
    return OK;
}

int tmi_impl(Workspace &w, const void *psi, bool f64, int n, int L, int block, int cycle, hipStream_t stream, double *out,
             double *terms, int order = 1)
{
    if (!psi || !out || L < 4 || L > n || block < 1 || 3 * block > L || cycle < 0 || cycle >= L)
    {
        set_error("tmi: bad arguments");
        return E_ARG;
    }
    auto range = [&](int begin, int cnt) {
        std::vector<int> v;
        for (int i = 0; i < cnt; ++i) v.push_back((begin + i) % L);
        return v;
    };
    const auto A = range(cycle, block), B = range(cycle + block, block), Cc = range(cycle + 2 * block, block);
    const auto D = range(cycle + 3 * block, L - 3 * block);
    auto cat = [](std::vector<int> x, const std::vector<int> &y) {
        x.insert(x.end(), y.begin(), y.end());
        return x;
    };
    const std::vector<int> sets[7] = {A, B, Cc, D, cat(A, B), cat(A, Cc), cat(B, Cc)};
    double s[7];
    for (int i = 0; i < 7; ++i)
        MIPT_HIP_TRY(entropy_impl(w, psi, f64, n, sets[i].data(), static_cast<int>(sets[i].size()), stream, order, &s[i]));
    *out = s[0] + s[1] + s[2] + s[3] - s[4] - s[5] - s[6];
    if (terms)
        for (int i = 0; i < 7; ++i) terms[i] = s[i];
    return OK;
}

int ensure_handle(Workspace &w, hipStream_t stream)
{
    if (!w.handle)
    {
        const rocblas_status s = rocblas_create_handle(&w.handle);
        if (s != rocblas_status_success)
        {
            set_error("rocblas_create_handle failed (%d)", static_cast<int>(s));
            return E_SOLVER + static_cast<int>(s);
        }
    }
    rocblas_set_stream(w.handle, stream);
    return OK;
}

// RDMs of `count` (state, kept set) items of one shape into dst (stride
// m*m), as the upper triangle of the row-major rho that the batched
// eigensolver reads: one gather and one Gram launch for all items.  roff and
// emask hold the uploaded tables; the caller keeps them until the stream
// has passed this work.
int build_rho_batch(Workspace &w, const float2 *const *psis, int n, const int *kept, int kept_count, int count, hipStream_t stream,
                    double2 *dst, std::vector<std::uint64_t> &roff, std::vector<std::uint64_t> &emask,
                    std::vector<std::uint64_t> &pv)
{
    const int m = 1 << kept_count, env = n - kept_count;
    const std::size_t K = std::size_t{1} << env;
    roff.assign(static_cast<std::size_t>(count) * m, 0);
    emask.assign(count, 0);
    pv.assign(count, 0);
    const std::uint64_t full = (n >= 64) ? ~std::uint64_t{0} : ((std::uint64_t{1} << n) - 1);
    for (int t = 0; t < count; ++t)
    {
        const int *kt = kept + static_cast<std::size_t>(t) * kept_count;
        std::uint64_t kmask;
        if (!psis[t] || validate_kept(n, kt, kept_count, kmask) != OK)
        {
            set_error("entropy batch: invalid state or kept-mode list (item %d)", t);
            return E_ARG;
        }
        emask[t] = full & ~kmask;
        pv[t] = reinterpret_cast<std::uint64_t>(psis[t]);
        for (int r = 0; r < m; ++r)
        {
            std::uint64_t o = 0;
            for (int j = 0; j < kept_count; ++j)
                if ((r >> j) & 1) o |= std::uint64_t{1} << kt[j];
            roff[static_cast<std::size_t>(t) * m + r] = o;
        }
    }
    MIPT_HIP_TRY(grow(w.rowb, w.rowb_cap, roff.size()));
    MIPT_HIP_TRY(grow(w.emb, w.emb_cap, emask.size()));
    MIPT_HIP_TRY(grow(w.psib, w.psib_cap, pv.size()));
    MIPT_HIP_CHECK(hipMemcpyAsync(w.rowb, roff.data(), roff.size() * 8, hipMemcpyHostToDevice, stream));
    MIPT_HIP_CHECK(hipMemcpyAsync(w.emb, emask.data(), emask.size() * 8, hipMemcpyHostToDevice, stream));
    MIPT_HIP_CHECK(hipMemcpyAsync(w.psib, pv.data(), pv.size() * 8, hipMemcpyHostToDevice, stream));
    const std::size_t plane = static_cast<std::size_t>(count) * m * K;
    MIPT_HIP_TRY(grow(w.ar, w.plane_cap, plane));
    MIPT_HIP_TRY(grow(w.ai, w.plane_cap_i, plane));
    const unsigned gblocks = static_cast<unsigned>(std::min<std::size_t>((plane + 255) / 256, 1u << 18));
    hipLaunchKernelGGL(gather_batch_kernel, dim3(gblocks), dim3(256), 0, stream,
                       reinterpret_cast<const float2 *const *>(w.psib), w.rowb, w.emb, kept_count, env, count, w.ar, w.ai);
    MIPT_HIP_CHECK(hipGetLastError());
    // RDMs of at most 32 rows take 32 x 32 tiles (a 64 x 64 tile would be
    // three quarters padding)
    const int ts = m <= 32 ? 32 : GT;
    const int nb = (m + ts - 1) / ts;
    const int ntri = nb * (nb + 1) / 2;
    int slot = -1;
    MIPT_HIP_TRY(gram_timing_begin(stream, 4.0 * m * (m + 1.0) * static_cast<double>(K) * count, slot));
    if (ts == 32)
        hipLaunchKernelGGL(gram_batch64_kernel<32>, dim3(static_cast<unsigned>(ntri), static_cast<unsigned>(count)),
                           dim3(256), 0, stream, w.ar, w.ai, m, static_cast<int>(K), ntri, dst);
    else
        hipLaunchKernelGGL(gram_batch64_kernel<GT>, dim3(static_cast<unsigned>(ntri), static_cast<unsigned>(count)),
                           dim3(256), 0, stream, w.ar, w.ai, m, static_cast<int>(K), ntri, dst);
    MIPT_HIP_CHECK(hipGetLastError());
    MIPT_HIP_TRY(gram_timing_end(stream, slot));
    return OK;
}

// Release batch-sized workspaces above 2 GiB so later allocations (other
// batches, other legs of a run) see the memory again.
int trim_batch_workspace(Workspace &w)
{
    const std::size_t held = w.rho_b_cap * 16 + (w.plane_cap + w.plane_cap_i) * 4 + w.eig_w_cap * 16;
    if (held <= (std::size_t{2} << 30)) return OK;
    MIPT_HIP_CHECK(hipFree(w.rho_b));
    MIPT_HIP_CHECK(hipFree(w.ar));
    MIPT_HIP_CHECK(hipFree(w.ai));
    MIPT_HIP_CHECK(hipFree(w.eig_w));
    w.rho_b = nullptr;
    w.ar = w.ai = nullptr;
    w.eig_w = nullptr;
    w.rho_b_cap = w.plane_cap = w.plane_cap_i = w.eig_w_cap = 0;
    return OK;
}

// entropy_many on the library's eigensolver: chunks of whole GPU rounds,
// each with one gather, one Gram, one tridiagonal reduction and one
// eigenvalue launch.
int entropy_many_own(Workspace &w, const float2 *const *psis, int n, const int *kept, int kept_count, int count, hipStream_t stream,
                     double *out)
{
    const int m = 1 << kept_count;
    const std::size_t K = std::size_t{1} << (n - kept_count);
    const std::size_t mm = static_cast<std::size_t>(m) * m;
    std::size_t free_b = 0, total_b = 0;
    MIPT_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    // per item: rho, the two FP32 planes, panel workspace, tridiagonal and result
    const std::size_t per = mm * 16 + static_cast<std::size_t>(m) * K * 8 + static_cast<std::size_t>(m) * 32 * 16 +
                            static_cast<std::size_t>(m) * 16 + 64;
    const std::size_t budget = std::min<std::size_t>(std::size_t{32} << 30, free_b / 4);
    const int maxchunk = static_cast<int>(std::max<std::size_t>(1, std::min<std::size_t>(count, budget / per)));
    const int nchunks = (count + maxchunk - 1) / maxchunk;
    const int chunk = (count + nchunks - 1) / nchunks;
    MIPT_HIP_TRY(grow(w.rho_b, w.rho_b_cap, static_cast<std::size_t>(chunk) * mm));
    std::vector<std::uint64_t> roff, emask, pv;
    for (int c0 = 0; c0 < count; c0 += chunk)
    {
        const int nb = std::min(chunk, count - c0);
        MIPT_HIP_TRY(build_rho_batch(w, psis + c0, n, kept + static_cast<std::size_t>(c0) * kept_count, kept_count, nb, stream,
                                     w.rho_b, roff, emask, pv));
        MIPT_HIP_TRY(own_entropies(w, m, nb, stream, w.rho_b, out + c0));  // synchronises the stream
    }
    return trim_batch_workspace(w);
}

// Von Neumann entropies (bits) of `count` items (state psis[i], kept set
// kept + i*kept_count), all of the same n and kept size: the RDMs are built
// into one strided buffer per chunk and their eigenvalues come from the
// library's batched eigensolver (one workgroup per matrix, eigen.hip) or,
// when selected or beyond EIG_MAX_N, from one
// rocsolver_zheevd_strided_batched call per chunk.
int entropy_many(Workspace &w, const float2 *const *psis, int n, const int *kept, int kept_count, int count, hipStream_t stream,
                 double *out)
{
    if (!psis || !out || count < 0 || (count > 0 && !kept))
    {
        set_error("entropy_many: null argument");
        return E_ARG;
    }
    if (kept_count == 0 || kept_count == n)
    {
        for (int i = 0; i < count; ++i) out[i] = 0.0;
        return OK;
    }
    const int m = 1 << kept_count;
    const std::size_t mm = static_cast<std::size_t>(m) * m;
    const bool own = g_eig_solver.load() == 0 && m <= EIG_MAX_N;
    if (own) return entropy_many_own(w, psis, n, kept, kept_count, count, stream, out);
    MIPT_HIP_TRY(ensure_handle(w, stream));
    if (m < 32)
    {
        // tiny RDMs on rocSOLVER: one solve each (cheap at this size)
        for (int i = 0; i < count; ++i)
        {
            int mo = 0;
            MIPT_HIP_TRY(build_rho(w, psis[i], false, n, kept + static_cast<std::size_t>(i) * kept_count, kept_count, stream, mo));
            std::vector<double> lam1;
            MIPT_HIP_TRY(eigvalsh(w, m, stream, lam1));
            out[i] = entropy_from_eigs(lam1);
        }
        return OK;
    }
    const int chunk = static_cast<int>(std::max<std::size_t>(1, std::min<std::size_t>(count, (std::size_t{1} << 30) / (mm * 16))));
    MIPT_HIP_TRY(grow(w.rho_b, w.rho_b_cap, chunk * mm));
    MIPT_HIP_TRY(grow(w.ev_b, w.ev_b_cap, static_cast<std::size_t>(chunk) * m));
    MIPT_HIP_TRY(grow(w.e_b, w.e_b_cap, static_cast<std::size_t>(chunk) * m));
    MIPT_HIP_TRY(grow(w.info_b, w.info_b_cap, static_cast<std::size_t>(chunk)));
    std::vector<double> lam(static_cast<std::size_t>(chunk) * m);
    std::vector<int> info(chunk);
    for (int c0 = 0; c0 < count; c0 += chunk)
    {
        const int nb = std::min(chunk, count - c0);
        for (int i = 0; i < nb; ++i)
        {
            int mo = 0;
            MIPT_HIP_TRY(build_rho(w, psis[c0 + i], false, n, kept + static_cast<std::size_t>(c0 + i) * kept_count,
                                   kept_count, stream, mo, w.rho_b + i * mm));
        }
        const rocblas_status st = rocsolver_zheevd_strided_batched(
            w.handle, rocblas_evect_none, rocblas_fill_lower, m, reinterpret_cast<rocblas_double_complex *>(w.rho_b), m,
            static_cast<rocblas_stride>(mm), w.ev_b, static_cast<rocblas_stride>(m), w.e_b,
            static_cast<rocblas_stride>(m), w.info_b, nb);
        if (st != rocblas_status_success)
        {
            set_error("rocsolver_zheevd_strided_batched failed (%d)", static_cast<int>(st));
            return E_SOLVER + static_cast<int>(st);
        }
        MIPT_HIP_CHECK(hipMemcpyAsync(lam.data(), w.ev_b, static_cast<std::size_t>(nb) * m * 8, hipMemcpyDeviceToHost,
                                      stream));
        MIPT_HIP_CHECK(hipMemcpyAsync(info.data(), w.info_b, nb * 4, hipMemcpyDeviceToHost, stream));
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        for (int i = 0; i < nb; ++i)
        {
            if (info[i] != 0)
            {
                set_error("FP64 zheevd did not converge (info=%d, item %d)", info[i], c0 + i);
                return E_EIG + info[i];
            }
            std::vector<double> li(lam.begin() + static_cast<std::size_t>(i) * m, lam.begin() + static_cast<std::size_t>(i + 1) * m);
            out[c0 + i] = entropy_from_eigs(li);
        }
    }
    return OK;
}

// Tr(rho^2) of `count` m x m matrices laid out every m*m entries, one block
// each, fixed reduction order.
__global__ void purity_many_kernel(const double2 *__restrict__ rho, int m, double *__restrict__ out)
{
    __shared__ double red[256];
    const std::size_t mm = static_cast<std::size_t>(m) * m;
    const double2 *r = rho + blockIdx.x * mm;
    double acc = 0.0;
    for (std::size_t i = threadIdx.x; i < mm; i += 256) acc += r[i].x * r[i].x + r[i].y * r[i].y;
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1)
    {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Purities Tr(rho^2) of `count` (state, kept set) items of one shape (the
// Renyi-2 counterpart of entropy_many: one strided RDM buffer per chunk).
int purity_many(Workspace &w, const float2 *const *psis, int n, const int *kept, int kept_count, int count, hipStream_t stream,
                double *out)
{
    if (!psis || !out || count < 0 || (count > 0 && !kept))
    {
        set_error("purity_many: null argument");
        return E_ARG;
    }
    if (kept_count == 0 || kept_count == n)
    {
        for (int i = 0; i < count; ++i) out[i] = 1.0;
        return OK;
    }
    const int m = 1 << kept_count;
    const std::size_t mm = static_cast<std::size_t>(m) * m;
    const int chunk = static_cast<int>(std::max<std::size_t>(1, std::min<std::size_t>(count, (std::size_t{1} << 30) / (mm * 16))));
    MIPT_HIP_TRY(grow(w.rho_b, w.rho_b_cap, chunk * mm));
    MIPT_HIP_TRY(grow(w.ev_b, w.ev_b_cap, static_cast<std::size_t>(chunk)));
    for (int c0 = 0; c0 < count; c0 += chunk)
    {
        const int nb = std::min(chunk, count - c0);
        for (int i = 0; i < nb; ++i)
        {
            int mo = 0;
            MIPT_HIP_TRY(build_rho(w, psis[c0 + i], false, n, kept + static_cast<std::size_t>(c0 + i) * kept_count,
                                   kept_count, stream, mo, w.rho_b + i * mm));
        }
        hipLaunchKernelGGL(purity_many_kernel, dim3(nb), dim3(256), 0, stream, w.rho_b, m, w.ev_b);
        MIPT_HIP_CHECK(hipGetLastError());
        MIPT_HIP_CHECK(hipMemcpyAsync(out + c0, w.ev_b, nb * sizeof(double), hipMemcpyDeviceToHost, stream));
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
    }
    return OK;
}

int rdm_export(Workspace &w, const void *psi, bool f64, int n, const int *kept, int count, double *host)
{
    if (!host || count > 12)
    {
        set_error("rdm_subsystem: bad arguments (kept <= 12)");
        return E_ARG;
    }
    int m = 0;
    if (count == 0)
    {
        host[0] = 1.0;
        host[1] = 0.0;
        return OK;
    }
    MIPT_HIP_TRY(build_rho(w, psi, f64, n, kept, count, nullptr, m));
    MIPT_HIP_CHECK(hipMemcpyAsync(host, w.rho, static_cast<std::size_t>(m) * m * 16, hipMemcpyDeviceToHost, nullptr));
    MIPT_HIP_CHECK(hipStreamSynchronize(nullptr));
    return OK;
}
} // namespace

int entropy_bits_f32(Workspace &w, const float2 *psi, int n, const int *kept, int kept_count, hipStream_t stream,
                     double *out)
{
    return entropy_impl(w, psi, false, n, kept, kept_count, stream, 1, out);
}

int tmi_bits_f32(Workspace &w, const float2 *psi, int n, int chain_sites, int block, int cycle, hipStream_t stream,
                 double *out, double *terms, int order)
{
    return tmi_impl(w, psi, false, n, chain_sites, block, cycle, stream, out, terms, order);
}

int renyi2_bits_f32(Workspace &w, const float2 *psi, int n, const int *kept, int kept_count, hipStream_t stream,
                    double *out)
{
    return entropy_impl(w, psi, false, n, kept, kept_count, stream, 2, out);
}

int entropy_many_bits_f32(Workspace &w, const float2 *const *psis, int n, const int *kept, int kept_count, int count,
                          hipStream_t stream, double *out)
{
    return entropy_many(w, psis, n, kept, kept_count, count, stream, out);
}

int purity_many_f32(Workspace &w, const float2 *const *psis, int n, const int *kept, int kept_count, int count,
                    hipStream_t stream, double *out)
{
    return purity_many(w, psis, n, kept, kept_count, count, stream, out);
}

int rho1_f32(Workspace &w, const float2 *psi, int n, int qubit, hipStream_t stream, double *out4)
{
    return rho1_impl<float2>(w, psi, n, qubit, stream, out4);
}

} // namespace mipt_hip

using namespace mipt_hip;

extern "C" {

static int check_fermion(int fermion_trace)
{
    if (fermion_trace != 0)
    {
        set_error("fermion_trace is not supported (qubit trace only)");
        return E_UNSUPPORTED;
    }
    return OK;
}

int mipt_hip_entropy_svd_f32(const void *s, int n, const int *k, int c, int f, double *out)
{
    MIPT_HIP_TRY(check_fermion(f));
    return entropy_impl(ws(), s, false, n, k, c, nullptr, 1, out);
}
int mipt_hip_entropy_svd_f64(const void *s, int n, const int *k, int c, int f, double *out)
{
    MIPT_HIP_TRY(check_fermion(f));
    return entropy_impl(ws(), s, true, n, k, c, nullptr, 1, out);
}
int mipt_hip_renyi2_f32(const void *s, int n, const int *k, int c, int f, double *out)
{
    MIPT_HIP_TRY(check_fermion(f));
    return entropy_impl(ws(), s, false, n, k, c, nullptr, 2, out);
}
int mipt_hip_renyi2_f64(const void *s, int n, const int *k, int c, int f, double *out)
{
    MIPT_HIP_TRY(check_fermion(f));
    return entropy_impl(ws(), s, true, n, k, c, nullptr, 2, out);
}
int mipt_hip_rdm_subsystem_f32(const void *s, int n, const int *k, int c, int f, double *rho)
{
    MIPT_HIP_TRY(check_fermion(f));
    return rdm_export(ws(), s, false, n, k, c, rho);
}
int mipt_hip_rdm_subsystem_f64(const void *s, int n, const int *k, int c, int f, double *rho)
{
    MIPT_HIP_TRY(check_fermion(f));
    return rdm_export(ws(), s, true, n, k, c, rho);
}
int mipt_hip_rho1_complex_f32(const void *s, int n, int q, double *rho)
{
    return rho1_impl<float2>(ws(), s, n, q, nullptr, rho);
}
int mipt_hip_rho1_complex_f64(const void *s, int n, int q, double *rho)
{
    return rho1_impl<double2>(ws(), s, n, q, nullptr, rho);
}
int mipt_hip_set_gram_profiling(int enable)
{
    GramTiming &g = gram_timing();
    std::lock_guard<std::mutex> lock(g.mu);
    for (auto &e : g.ev)
    {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    g.ev.clear();
    g.flops.clear();
    g.on = enable != 0;
    return OK;
}

int mipt_hip_gram_stats(double *avg_ms, int *count, double *flops_avg)
{
    GramTiming &g = gram_timing();
    std::lock_guard<std::mutex> lock(g.mu);
    double ms_sum = 0.0, fl_sum = 0.0;
    for (std::size_t i = 0; i < g.ev.size(); ++i)
    {
        MIPT_HIP_CHECK(hipEventSynchronize(g.ev[i].second));
        float ms = 0.f;
        MIPT_HIP_CHECK(hipEventElapsedTime(&ms, g.ev[i].first, g.ev[i].second));
        ms_sum += ms;
        fl_sum += g.flops[i];
    }
    const double n = static_cast<double>(g.ev.size());
    if (avg_ms) *avg_ms = n > 0 ? ms_sum / n : 0.0;
    if (count) *count = static_cast<int>(g.ev.size());
    if (flops_avg) *flops_avg = n > 0 ? fl_sum / n : 0.0;
    return OK;
}

int mipt_hip_set_gram_precision(int bits)
{
    if (bits != 32 && bits != 64)
    {
        set_error("gram precision must be 32 or 64, got %d", bits);
        return E_ARG;
    }
    g_gram_bits.store(bits);
    return OK;
}

int mipt_hip_entropy_hermitian_batched(const double *rho_ri, int m, int count, double *entropy_bits,
                                       double *eigenvalues)
{
    if (!rho_ri || !entropy_bits || m < 1 || m > EIG_MAX_N || count < 0)
    {
        set_error("entropy_hermitian_batched: bad arguments (m=%d in [1, %d], count=%d)", m, EIG_MAX_N, count);
        return E_ARG;
    }
    if (count == 0) return OK;
    Workspace &w = ws();
    hipStream_t stream = nullptr;
    const std::size_t mm = static_cast<std::size_t>(m) * m, total = mm * count;
    MIPT_HIP_TRY(grow(w.rho_b, w.rho_b_cap, total));
    MIPT_HIP_TRY(grow(w.eig_w, w.eig_w_cap, static_cast<std::size_t>(count) * m * 32));
    MIPT_HIP_TRY(grow(w.eig_d, w.eig_d_cap, static_cast<std::size_t>(count) * m));
    MIPT_HIP_TRY(grow(w.eig_e, w.eig_e_cap, static_cast<std::size_t>(count) * m));
    MIPT_HIP_TRY(grow(w.eig_s, w.eig_s_cap, static_cast<std::size_t>(count)));
    double *lam = nullptr;
    if (eigenvalues) MIPT_HIP_TRY(grow(w.ev_b, w.ev_b_cap, static_cast<std::size_t>(count) * m));
    if (eigenvalues) lam = w.ev_b;
    // row-major rho read as column-major is rho^T = conj(rho): the same spectrum
    MIPT_HIP_CHECK(hipMemcpyAsync(w.rho_b, rho_ri, total * sizeof(double2), hipMemcpyHostToDevice, stream));
    MIPT_HIP_TRY(launch_hermitian_entropies(w.rho_b, m, mm, count, w.eig_w, w.eig_d, w.eig_e, w.eig_s, lam, stream));
    MIPT_HIP_CHECK(hipMemcpyAsync(entropy_bits, w.eig_s, static_cast<std::size_t>(count) * sizeof(double),
                                  hipMemcpyDeviceToHost, stream));
    if (eigenvalues)
        MIPT_HIP_CHECK(hipMemcpyAsync(eigenvalues, lam, static_cast<std::size_t>(count) * m * sizeof(double),
                                      hipMemcpyDeviceToHost, stream));
    MIPT_HIP_CHECK(hipStreamSynchronize(stream));
    return OK;
}

int mipt_hip_set_eigensolver(int which)
{
    if (which != 0 && which != 1)
    {
        set_error("eigensolver must be 0 (library kernels) or 1 (rocSOLVER), got %d", which);
        return E_ARG;
    }
    g_eig_solver.store(which);
    return OK;
}

int mipt_hip_tmi_f32(const void *s, int n, int L, int block, int cycle, double *out, double *terms)
{
    return tmi_impl(ws(), s, false, n, L, block, cycle, nullptr, out, terms);
}

int mipt_hip_tmi_f64(const void *s, int n, int L, int block, int cycle, double *out, double *terms)
{
    return tmi_impl(ws(), s, true, n, L, block, cycle, nullptr, out, terms);
}

} // extern "C"
