// State-vector kernels shared by the plain engine and the batched engine's
// overflow path: dense k-target apply, deterministic FP64 conditional
// histograms over measured bits, and collapse.  Real = float (complex64
// state, FP32 arithmetic) or double (complex128).
#pragma once

#include "common.hpp"

namespace mipt_hip
{

template <typename Real>
struct CxT;
template <>
struct CxT<float>
{
    using type = float2;
};
template <>
struct CxT<double>
{
    using type = double2;
};

constexpr int HIST_THREADS = 256;
constexpr int HIST_PER_THREAD = 16;

// Dense 2^K x 2^K row-major matrix on K targets (bit j of the local index =
// targets[j]); one thread per group of 2^K amplitudes.  `sorted` holds the
// targets in ascending order for the zero-bit insertion.
struct DenseTargets
{
    int k;
    int sorted[8];
    std::uint64_t off[256];
};

template <typename Real, int K>
__global__ void __launch_bounds__(256) apply_dense_kernel(typename CxT<Real>::type *__restrict__ psi, std::uint64_t groups,
                                                          const DenseTargets *__restrict__ tg,
                                                          const typename CxT<Real>::type *__restrict__ mat)
{
    using C = typename CxT<Real>::type;
    constexpr int D = 1 << K;
    __shared__ C m[D * D];
    __shared__ std::uint64_t off[D];
    __shared__ int sorted[K];
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) m[i] = mat[i];
    for (int i = threadIdx.x; i < D; i += blockDim.x) off[i] = tg->off[i];
    if (threadIdx.x < K) sorted[threadIdx.x] = tg->sorted[threadIdx.x];
    __syncthreads();
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
    for (std::uint64_t g = blockIdx.x * static_cast<std::uint64_t>(blockDim.x) + threadIdx.x; g < groups; g += stride)
    {
        std::uint64_t base = g;
#pragma unroll
        for (int j = 0; j < K; ++j)
        {
            const int t = sorted[j];
            base = ((base >> t) << (t + 1)) | (base & ((std::uint64_t{1} << t) - 1));
        }
        C x[D];
#pragma unroll
        for (int a = 0; a < D; ++a) x[a] = psi[base | off[a]];
#pragma unroll
        for (int r = 0; r < D; ++r)
        {
            Real yr = 0, yi = 0;
#pragma unroll
            for (int c = 0; c < D; ++c)
            {
                const C u = m[r * D + c];
                yr = fma(u.x, x[c].x, yr);
                yr = fma(-u.y, x[c].y, yr);
                yi = fma(u.x, x[c].y, yi);
                yi = fma(u.y, x[c].x, yi);
            }
            psi[base | off[r]] = C{yr, yi};
        }
    }
}

// Conditional FP64 histogram, stage 1.  bin b over `mmask` (bit j of b =
// j-th set bit of mmask), restricted to amplitudes whose `cmask` bits equal
// `cval`.  Rest index r enumerates the bits outside mmask; each thread sums
// HIST_PER_THREAD rest values in order, each block reduces in a fixed tree
// and writes partial[traj][bin][blk].  Deterministic.
template <typename Real>
__global__ void __launch_bounds__(HIST_THREADS) hist_partial_kernel(
    const typename CxT<Real>::type *__restrict__ states, const std::uint64_t *__restrict__ state_off, int n,
    const std::uint64_t *__restrict__ mmask_v, const std::uint64_t *__restrict__ cmask_v,
    const std::uint64_t *__restrict__ cval_v, int blocks_per_bin, double *__restrict__ partial, int partial_stride)
{
    const int traj = blockIdx.y;
    const typename CxT<Real>::type *psi = states + state_off[traj];
    const std::uint64_t mmask = mmask_v[traj], cmask = cmask_v[traj], cval = cval_v[traj];
    const std::uint64_t dim_mask = (n >= 64) ? ~std::uint64_t{0} : ((std::uint64_t{1} << n) - 1);
    const std::uint64_t rest_mask = dim_mask & ~mmask;
    const int m = __popcll(mmask);
    const std::uint64_t rest_count = std::uint64_t{1} << (n - m);
    const std::uint64_t bins = std::uint64_t{1} << m;
    const std::uint64_t bin = blockIdx.x / blocks_per_bin;
    const int blk = blockIdx.x % blocks_per_bin;
    double acc = 0.0;
    if (bin < bins)
    {
        const std::uint64_t bin_bits = pdep64(bin, mmask);
        const std::uint64_t first = (static_cast<std::uint64_t>(blk) * HIST_THREADS + threadIdx.x) * HIST_PER_THREAD;
        if (first < rest_count)
        {
            std::uint64_t r = pdep64(first, rest_mask);
            for (int e = 0; e < HIST_PER_THREAD && first + e < rest_count; ++e)
            {
                const std::uint64_t i = bin_bits | r;
                if ((i & cmask) == cval)
                {
                    const auto v = psi[i];
                    acc = fma(static_cast<double>(v.x), static_cast<double>(v.x), acc);
                    acc = fma(static_cast<double>(v.y), static_cast<double>(v.y), acc);
                }
                r = ((r | ~rest_mask) + 1) & rest_mask;
            }
        }
    }
    __shared__ double red[HIST_THREADS];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = HIST_THREADS / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && bin < bins)
        partial[static_cast<std::size_t>(traj) * partial_stride + bin * blocks_per_bin + blk] = red[0];
}

// Stage 2: hist[traj][bin] = sum over blk in order.
static __global__ void hist_final_kernel(const double *__restrict__ partial, int partial_stride, int blocks_per_bin,
                                  const std::uint64_t *__restrict__ mmask_v, double *__restrict__ hist, int hist_stride)
{
    const int traj = blockIdx.y;
    const std::uint64_t bins = std::uint64_t{1} << __popcll(mmask_v[traj]);
    const std::uint64_t bin = blockIdx.x * static_cast<std::uint64_t>(blockDim.x) + threadIdx.x;
    if (bin >= bins) return;
    const double *p = partial + static_cast<std::size_t>(traj) * partial_stride + bin * blocks_per_bin;
    double s = 0.0;
    for (int b = 0; b < blocks_per_bin; ++b) s += p[b];
    hist[static_cast<std::size_t>(traj) * hist_stride + bin] = s;
}

// psi[i] <- (i & cmask) == cval ? psi[i] * scale : 0
template <typename Real>
__global__ void collapse_kernel(typename CxT<Real>::type *__restrict__ psi, std::uint64_t dim, std::uint64_t cmask,
                                std::uint64_t cval, Real scale)
{
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
    for (std::uint64_t i = blockIdx.x * static_cast<std::uint64_t>(blockDim.x) + threadIdx.x; i < dim; i += stride)
    {
        auto v = psi[i];
        if ((i & cmask) == cval)
        {
            v.x *= scale;
            v.y *= scale;
        }
        else
        {
            v.x = 0;
            v.y = 0;
        }
        psi[i] = v;
    }
}

template <typename Real>
__global__ void set_basis_kernel(typename CxT<Real>::type *__restrict__ psi, std::uint64_t dim, std::uint64_t index)
{
    const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
    for (std::uint64_t i = blockIdx.x * static_cast<std::uint64_t>(blockDim.x) + threadIdx.x; i < dim; i += stride)
    {
        typename CxT<Real>::type v;
        v.x = (i == index) ? Real(1) : Real(0);
        v.y = 0;
        psi[i] = v;
    }
}

// Sequential Born decisions from a joint histogram over `m` ascending sites
// (bin bit j <-> site j of the chunk).  Given the outcomes already fixed for
// earlier chunks (folded into the histogram as a condition), site j's
// conditional P(1) is W(prefix, 1) / W(prefix) and its outcome is 1 iff
// u_j < P(1) (cpp/include/mipt/free_energy_measure.hpp:59).  Returns the
// chosen bin; p1[j] receives each conditional P(1); the weight of the chosen
// bin is the raw joint branch weight used for renormalisation.
__host__ __device__ inline std::uint32_t sequential_from_hist(const double *h, int m, const double *u, double *p1,
                                                              int *status)
{
    std::uint32_t prefix = 0;
    *status = 0;
    for (int j = 0; j < m; ++j)
    {
        const std::uint32_t pmask = (1u << j) - 1u;
        double w = 0.0, w1 = 0.0;
        for (std::uint32_t b = 0; b < (1u << m); ++b)
        {
            if ((b & pmask) != prefix) continue;
            w += h[b];
            if ((b >> j) & 1u) w1 += h[b];
        }
        if (!(w > 0.0))
        {
            *status = E_NORM;
            return prefix;
        }
        double p = w1 / w;
        p = p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);
        p1[j] = p;
        if (u[j] < p) prefix |= (1u << j);
    }
    return prefix;
}

} // namespace mipt_hip
