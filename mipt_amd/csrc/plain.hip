// Plain per-trajectory engine (one launch per blocked op, a host round trip
// per measured layer) and the sequential-measurement drop-in.  Structure and
// semantics of cusv::Engine (cpp/include/mipt/cusv/engine.hpp:100-582) and of
// mipt_cuda_measure_collapse_* (cpp/src/cuda/free_energy_measure.cu:127-299),
// written for gfx950 with deterministic FP64 histograms.
#include "mipt_hip.h"
#include "state_ops.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

namespace mipt_hip
{
static thread_local char g_error[512] = {0};

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

const char *last_error() { return g_error; }

namespace
{
int grid_for(std::uint64_t work, int threads)
{
    const std::uint64_t b = (work + threads - 1) / threads;
    return static_cast<int>(b < 65536 ? (b ? b : 1) : 65536);
}

// Deterministic conditional histogram of one state on the engine's stream.
struct HistScratch
{
    double *partial = nullptr;
    std::size_t partial_cap = 0;
    double *hist = nullptr;
    std::size_t hist_cap = 0;
    std::uint64_t *masks = nullptr;  // off, mmask, cmask, cval
    ~HistScratch()
    {
        if (partial) (void)hipFree(partial);
        if (hist) (void)hipFree(hist);
        if (masks) (void)hipFree(masks);
    }
};

template <typename Real>
int conditional_hist(const void *psi, int n, std::uint64_t mmask, std::uint64_t cmask, std::uint64_t cval,
                     hipStream_t stream, HistScratch &s, std::vector<double> &out)
{
    const int m = __builtin_popcountll(mmask);
    const std::uint64_t bins = std::uint64_t{1} << m;
    const std::uint64_t rest = std::uint64_t{1} << (n - m);
    const std::uint64_t per_block = static_cast<std::uint64_t>(HIST_THREADS) * HIST_PER_THREAD;
    const int bpb = static_cast<int>((rest + per_block - 1) / per_block);
    const std::size_t need_partial = static_cast<std::size_t>(bins) * bpb;
    if (need_partial > s.partial_cap)
    {
        if (s.partial) MIPT_HIP_CHECK(hipFree(s.partial));
        s.partial = nullptr;
        MIPT_HIP_CHECK(hipMalloc(&s.partial, need_partial * sizeof(double)));
        s.partial_cap = need_partial;
    }
    if (bins > s.hist_cap)
    {
        if (s.hist) MIPT_HIP_CHECK(hipFree(s.hist));
        s.hist = nullptr;
        MIPT_HIP_CHECK(hipMalloc(&s.hist, bins * sizeof(double)));
        s.hist_cap = bins;
    }
    if (!s.masks) MIPT_HIP_CHECK(hipMalloc(&s.masks, 4 * sizeof(std::uint64_t)));
    const std::uint64_t hm[4] = {0, mmask, cmask, cval};
    MIPT_HIP_CHECK(hipMemcpyAsync(s.masks, hm, sizeof(hm), hipMemcpyHostToDevice, stream));
    const dim3 g1(static_cast<unsigned>(bins * bpb), 1);
    hipLaunchKernelGGL(hist_partial_kernel<Real>, g1, dim3(HIST_THREADS), 0, stream,
                       static_cast<const typename CxT<Real>::type *>(psi), s.masks + 0, n, s.masks + 1, s.masks + 2,
                       s.masks + 3, bpb, s.partial, static_cast<int>(need_partial));
    MIPT_HIP_CHECK(hipGetLastError());
    const dim3 g2(static_cast<unsigned>((bins + 255) / 256), 1);
    hipLaunchKernelGGL(hist_final_kernel, g2, dim3(256), 0, stream, s.partial, static_cast<int>(need_partial), bpb,
                       s.masks + 1, s.hist, static_cast<int>(bins));
    MIPT_HIP_CHECK(hipGetLastError());
    out.resize(bins);
    MIPT_HIP_CHECK(hipMemcpyAsync(out.data(), s.hist, bins * sizeof(double), hipMemcpyDeviceToHost, stream));
    MIPT_HIP_CHECK(hipStreamSynchronize(stream));
    return OK;
}

template <typename Real>
int collapse(void *psi, int n, std::uint64_t cmask, std::uint64_t cval, double scale, hipStream_t stream)
{
    const std::uint64_t dim = std::uint64_t{1} << n;
    hipLaunchKernelGGL(collapse_kernel<Real>, dim3(grid_for(dim, 256)), dim3(256), 0, stream,
                       static_cast<typename CxT<Real>::type *>(psi), dim, cmask, cval, static_cast<Real>(scale));
    MIPT_HIP_CHECK(hipGetLastError());
    return OK;
}

// Never freed (like the reference's thread_local device buffers): freeing at
// thread exit could run after the HIP runtime is torn down.
thread_local HistScratch *t_measure_scratch = nullptr;
thread_local char t_measure_error[512] = {0};

template <typename Real>
int measure_collapse_impl(void *state, int n, int q, double u, int *outcome, double *prob)
{
    t_measure_error[0] = '\0';
    if (state == nullptr || outcome == nullptr || prob == nullptr || n < 1 || n >= 63 || q < 0 || q >= n ||
        !(u >= 0.0 && u < 1.0))
    {
        std::snprintf(t_measure_error, sizeof(t_measure_error), "Invalid sequential-measurement argument.");
        set_error("%s", t_measure_error);
        return E_ARG;
    }
    if (!t_measure_scratch) t_measure_scratch = new HistScratch();
    std::vector<double> h;
    int st = conditional_hist<Real>(state, n, std::uint64_t{1} << q, 0, 0, nullptr, *t_measure_scratch, h);
    if (st != 0)
    {
        std::snprintf(t_measure_error, sizeof(t_measure_error), "%s", last_error());
        return st;
    }
    const double total = h[0] + h[1];
    const double one = h[1];
    if (!(total > 0.0) || !std::isfinite(total) || !std::isfinite(one) || one < 0.0 || one > total * (1.0 + 1.0e-10))
    {
        std::snprintf(t_measure_error, sizeof(t_measure_error), "Invalid branch norms: total=%.17g, one=%.17g",
                      total, one);
        set_error("%s", t_measure_error);
        return E_NORM;
    }
    const double p1 = std::fmin(1.0, std::fmax(0.0, one / total));
    *outcome = u < p1 ? 1 : 0;
    *prob = *outcome ? p1 : 1.0 - p1;
    const double branch = *outcome ? one : total - one;
    if (!(*prob > 0.0) || !std::isfinite(*prob))
    {
        std::snprintf(t_measure_error, sizeof(t_measure_error),
                      "Sampled an outcome with invalid conditional probability %.17g", *prob);
        set_error("%s", t_measure_error);
        return E_PROB;
    }
    if (!(branch > 0.0) || !std::isfinite(branch))
    {
        std::snprintf(t_measure_error, sizeof(t_measure_error), "Sampled an outcome with invalid branch norm %.17g",
                      branch);
        set_error("%s", t_measure_error);
        return E_BRANCH;
    }
    const std::uint64_t mask = std::uint64_t{1} << q;
    st = collapse<Real>(state, n, mask, *outcome ? mask : 0, 1.0 / std::sqrt(branch), nullptr);
    if (st == 0)
    {
        const hipError_t e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess)
        {
            set_error("collapse synchronization: %s", hipGetErrorString(e));
            st = static_cast<int>(e);
        }
    }
    if (st != 0) std::snprintf(t_measure_error, sizeof(t_measure_error), "%s", last_error());
    return st;
}

// ---------------------------------------------------------- allocation-free
// The sequential measurement for callers that cannot allocate device memory.
// Its reduction partials, histogram and masks live in `scratch`, file-scope
// device memory of a fixed size (one copy per GPU, held until the process
// ends): 4 mask words, 2 histogram bins, 2 spare words, then the partials of
// the stage-1 reduction, 2^(n-12) doubles for n >= 13 -- at most 2^17 (1 MiB)
// for n <= 29.  Every word a call reads is written earlier in the same call,
// so nothing needs clearing; calls that overlap on one GPU race on it.
constexpr int NOALLOC_MAX_QUBITS = 29;
constexpr std::size_t NOALLOC_WORDS = 8 + (std::size_t{1} << (NOALLOC_MAX_QUBITS - 12));
static __device__ __align__(256) double scratch[NOALLOC_WORDS];

// The device address of `scratch` on each GPU, looked up once per device with
// that device current (the array's host-side name is not a device address).
struct NoAllocScratch
{
    std::mutex mu;
    double *ptr[64] = {};
};
NoAllocScratch &noalloc_scratch()
{
    static NoAllocScratch s;
    return s;
}

int noalloc_scratch_ptr(double **out)
{
    int dev = 0;
    MIPT_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64)
    {
        set_error("allocation-free measurement: device %d out of range", dev);
        return E_ARG;
    }
    NoAllocScratch &s = noalloc_scratch();
    std::lock_guard<std::mutex> lock(s.mu);
    if (!s.ptr[dev])
    {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
        {
            set_error("allocation-free measurement: scratch address on device %d: %s", dev, hipGetErrorString(e));
            return e != hipSuccess ? static_cast<int>(e) : E_ARG;
        }
        s.ptr[dev] = static_cast<double *>(p);
    }
    *out = s.ptr[dev];
    return OK;
}

__global__ void set_masks_kernel(std::uint64_t *m, std::uint64_t off, std::uint64_t mmask, std::uint64_t cmask,
                                 std::uint64_t cval)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
    {
        m[0] = off;
        m[1] = mmask;
        m[2] = cmask;
        m[3] = cval;
    }
}

// measure_collapse_impl with its buffers in `scratch`: the same kernels, the
// same fixed-order FP64 sums and the same decision, so the same outcome,
// probability and collapsed state.
template <typename Real>
int measure_collapse_noalloc_impl(void *state, int n, int q, double u, int *outcome, double *prob)
{
    t_measure_error[0] = '\0';
    if (state == nullptr || outcome == nullptr || prob == nullptr || n < 1 || n >= 63 || q < 0 || q >= n ||
        !(u >= 0.0 && u < 1.0))
    {
        std::snprintf(t_measure_error, sizeof(t_measure_error), "Invalid sequential-measurement argument.");
        set_error("%s", t_measure_error);
        return E_ARG;
    }
    if (n > NOALLOC_MAX_QUBITS)
    {
        std::snprintf(t_measure_error, sizeof(t_measure_error),
                      "allocation-free measurement: %d qubits exceed its fixed scratch (at most %d qubits)", n,
                      NOALLOC_MAX_QUBITS);
        set_error("%s", t_measure_error);
        return E_UNSUPPORTED;
    }
    double *base = nullptr;
    int st = noalloc_scratch_ptr(&base);
    if (st != OK)
    {
        std::snprintf(t_measure_error, sizeof(t_measure_error), "%s", last_error());
        return st;
    }
    std::uint64_t *masks = reinterpret_cast<std::uint64_t *>(base);
    double *hist = base + 4;
    double *partial = base + 8;
    const std::uint64_t mmask = std::uint64_t{1} << q;
    const std::uint64_t rest = std::uint64_t{1} << (n - 1);
    const std::uint64_t per_block = static_cast<std::uint64_t>(HIST_THREADS) * HIST_PER_THREAD;
    const int bpb = static_cast<int>((rest + per_block - 1) / per_block);
    const std::size_t need_partial = std::size_t{2} * bpb;  // within NOALLOC_WORDS - 8 for n <= 29
    hipLaunchKernelGGL(set_masks_kernel, dim3(1), dim3(64), 0, nullptr, masks, std::uint64_t{0}, mmask,
                       std::uint64_t{0}, std::uint64_t{0});
    hipLaunchKernelGGL(hist_partial_kernel<Real>, dim3(static_cast<unsigned>(2 * bpb), 1), dim3(HIST_THREADS), 0, nullptr,
                       static_cast<const typename CxT<Real>::type *>(state), masks + 0, n, masks + 1, masks + 2,
                       masks + 3, bpb, partial, static_cast<int>(need_partial));
    hipLaunchKernelGGL(hist_final_kernel, dim3(1, 1), dim3(256), 0, nullptr, partial, static_cast<int>(need_partial), bpb,
                       masks + 1, hist, 2);
    double h[2] = {0.0, 0.0};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess)
    {
        set_error("allocation-free measurement: %s", hipGetErrorString(e));
        std::snprintf(t_measure_error, sizeof(t_measure_error), "%s", last_error());
        return static_cast<int>(e);
    }
    const double total = h[0] + h[1];
    const double one = h[1];
    if (!(total > 0.0) || !std::isfinite(total) || !std::isfinite(one) || one < 0.0 || one > total * (1.0 + 1.0e-10))
    {
        std::snprintf(t_measure_error, sizeof(t_measure_error), "Invalid branch norms: total=%.17g, one=%.17g",
                      total, one);
        set_error("%s", t_measure_error);
        return E_NORM;
    }
    const double p1 = std::fmin(1.0, std::fmax(0.0, one / total));
    *outcome = u < p1 ? 1 : 0;
    *prob = *outcome ? p1 : 1.0 - p1;
    const double branch = *outcome ? one : total - one;
    if (!(*prob > 0.0) || !std::isfinite(*prob))
    {
        std::snprintf(t_measure_error, sizeof(t_measure_error),
                      "Sampled an outcome with invalid conditional probability %.17g", *prob);
        set_error("%s", t_measure_error);
        return E_PROB;
    }
    if (!(branch > 0.0) || !std::isfinite(branch))
    {
        std::snprintf(t_measure_error, sizeof(t_measure_error), "Sampled an outcome with invalid branch norm %.17g",
                      branch);
        set_error("%s", t_measure_error);
        return E_BRANCH;
    }
    st = collapse<Real>(state, n, mmask, *outcome ? mmask : 0, 1.0 / std::sqrt(branch), nullptr);
    if (st == 0)
    {
        e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess)
        {
            set_error("collapse synchronization: %s", hipGetErrorString(e));
            st = static_cast<int>(e);
        }
    }
    if (st != 0) std::snprintf(t_measure_error, sizeof(t_measure_error), "%s", last_error());
    return st;
}
} // namespace

// ------------------------------------------------------------------ engine
struct PlainEngine
{
    bool fp64 = false;
    int n = 0;
    void *state = nullptr;
    void *owned = nullptr;
    hipStream_t stream = nullptr;
    void *mat_dev = nullptr;
    DenseTargets *tg_dev = nullptr;
    HistScratch hs;
    std::vector<double> hist;
    // one trajectory's blocked matrices and target tables, staged at once
    void *traj_mat = nullptr;
    std::size_t traj_mat_cap = 0;
    DenseTargets *traj_tg = nullptr;
    std::size_t traj_tg_cap = 0;

    ~PlainEngine()
    {
        if (owned) (void)hipFree(owned);
        if (mat_dev) (void)hipFree(mat_dev);
        if (tg_dev) (void)hipFree(tg_dev);
        if (traj_mat) (void)hipFree(traj_mat);
        if (traj_tg) (void)hipFree(traj_tg);
    }

    std::size_t elem() const { return fp64 ? 16 : 8; }
    std::size_t dim() const { return std::size_t{1} << n; }

    int ensure_scratch()
    {
        if (!mat_dev) MIPT_HIP_CHECK(hipMalloc(&mat_dev, 256 * 256 * 16));
        if (!tg_dev) MIPT_HIP_CHECK(hipMalloc(&tg_dev, sizeof(DenseTargets)));
        return OK;
    }

    template <typename Real>
    int launch_dense(int k, std::uint64_t groups)
    {
        using C = typename CxT<Real>::type;
        auto *psi = static_cast<C *>(state);
        const auto *m = static_cast<const C *>(mat_dev);
        const dim3 grid(grid_for(groups, 256));
        switch (k)
        {
        case 1: hipLaunchKernelGGL((apply_dense_kernel<Real, 1>), grid, dim3(256), 0, stream, psi, groups, tg_dev, m); break;
        case 2: hipLaunchKernelGGL((apply_dense_kernel<Real, 2>), grid, dim3(256), 0, stream, psi, groups, tg_dev, m); break;
        case 3: hipLaunchKernelGGL((apply_dense_kernel<Real, 3>), grid, dim3(256), 0, stream, psi, groups, tg_dev, m); break;
        case 4: hipLaunchKernelGGL((apply_dense_kernel<Real, 4>), grid, dim3(256), 0, stream, psi, groups, tg_dev, m); break;
        case 5: hipLaunchKernelGGL((apply_dense_kernel<Real, 5>), grid, dim3(256), 0, stream, psi, groups, tg_dev, m); break;
        default: set_error("dense apply supports 1..5 targets, got %d", k); return E_UNSUPPORTED;
        }
        MIPT_HIP_CHECK(hipGetLastError());
        return OK;
    }

    // H2D of the matrix and targets per call, as the reference engine does.
    int apply_matrix(const int *targets, int k, const double *m_ri)
    {
        if (!state || k < 1 || k > 5 || k > n) { set_error("apply_matrix: bad arguments"); return E_ARG; }
        MIPT_HIP_TRY(ensure_scratch());
        DenseTargets tg{};
        tg.k = k;
        std::uint64_t tmask = 0;
        for (int j = 0; j < k; ++j)
        {
            if (targets[j] < 0 || targets[j] >= n || ((tmask >> targets[j]) & 1)) { set_error("apply_matrix: bad target"); return E_ARG; }
            tmask |= std::uint64_t{1} << targets[j];
            tg.sorted[j] = targets[j];
        }
        std::sort(tg.sorted, tg.sorted + k);
        const int d = 1 << k;
        for (int a = 0; a < d; ++a)
            for (int j = 0; j < k; ++j)
                if ((a >> j) & 1) tg.off[a] |= std::uint64_t{1} << targets[j];
        const std::size_t cnt = static_cast<std::size_t>(d) * d;
        MIPT_HIP_CHECK(hipMemcpyAsync(tg_dev, &tg, sizeof(tg), hipMemcpyHostToDevice, stream));
        if (fp64)
        {
            MIPT_HIP_CHECK(hipMemcpyAsync(mat_dev, m_ri, cnt * 16, hipMemcpyHostToDevice, stream));
            MIPT_HIP_TRY(launch_dense<double>(k, dim() >> k));
        }
        else
        {
            std::vector<float> narrow(2 * cnt);
            for (std::size_t i = 0; i < 2 * cnt; ++i) narrow[i] = static_cast<float>(m_ri[i]);
            MIPT_HIP_CHECK(hipMemcpyAsync(mat_dev, narrow.data(), cnt * 8, hipMemcpyHostToDevice, stream));
            MIPT_HIP_TRY(launch_dense<float>(k, dim() >> k));
        }
        // The host staging buffers must outlive the async copies.
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        return OK;
    }

    // Kronecker blocking of consecutive disjoint bonds up to block_targets
    // (cpp/include/mipt/cusv/gates.hpp:295-333; tensor order of
    // cpp/include/mipt/cusv/matrix.hpp:214-235: later ops take the high bits).
    // Calls emit(targets, k, matrix) once per block, in order.
    template <typename Emit>
    static int block_bonds(const int *q0, const int *q1, const double *gates, int count, int block_targets, Emit emit)
    {
        int i = 0;
        while (i < count)
        {
            std::vector<int> targets = {q0[i], q1[i]};
            std::vector<double> m(gates + 32 * i, gates + 32 * (i + 1));
            int j = i + 1;
            while (j < count && static_cast<int>(targets.size()) + 2 <= block_targets)
            {
                if (std::find(targets.begin(), targets.end(), q0[j]) != targets.end() ||
                    std::find(targets.begin(), targets.end(), q1[j]) != targets.end())
                    break;
                const int dl = 1 << targets.size();
                std::vector<double> out(static_cast<std::size_t>(dl * 4) * (dl * 4) * 2, 0.0);
                const double *hi = gates + 32 * j;
                for (int hr = 0; hr < 4; ++hr)
                    for (int hc = 0; hc < 4; ++hc)
                    {
                        const double hvr = hi[2 * (hr * 4 + hc)], hvi = hi[2 * (hr * 4 + hc) + 1];
                        for (int lr = 0; lr < dl; ++lr)
                            for (int lc = 0; lc < dl; ++lc)
                            {
                                const double lvr = m[2 * (lr * dl + lc)], lvi = m[2 * (lr * dl + lc) + 1];
                                const std::size_t o = 2 * (static_cast<std::size_t>(hr * dl + lr) * (dl * 4) + (hc * dl + lc));
                                out[o] = hvr * lvr - hvi * lvi;
                                out[o + 1] = hvr * lvi + hvi * lvr;
                            }
                    }
                m.swap(out);
                targets.push_back(q0[j]);
                targets.push_back(q1[j]);
                ++j;
            }
            MIPT_HIP_TRY(emit(targets.data(), static_cast<int>(targets.size()), m.data()));
            i = j;
        }
        return OK;
    }

    int apply_bonds(const int *q0, const int *q1, const double *gates, int count, int block_targets)
    {
        return block_bonds(q0, q1, gates, count, block_targets,
                           [&](const int *t, int k, const double *m) { return apply_matrix(t, k, m); });
    }

    template <typename Real>
    int launch_dense_at(int k, std::uint64_t groups, const DenseTargets *tg, const void *mat)
    {
        using C = typename CxT<Real>::type;
        auto *psi = static_cast<C *>(state);
        const auto *m = static_cast<const C *>(mat);
        const dim3 grid(grid_for(groups, 256));
        switch (k)
        {
        case 1: hipLaunchKernelGGL((apply_dense_kernel<Real, 1>), grid, dim3(256), 0, stream, psi, groups, tg, m); break;
        case 2: hipLaunchKernelGGL((apply_dense_kernel<Real, 2>), grid, dim3(256), 0, stream, psi, groups, tg, m); break;
        case 3: hipLaunchKernelGGL((apply_dense_kernel<Real, 3>), grid, dim3(256), 0, stream, psi, groups, tg, m); break;
        case 4: hipLaunchKernelGGL((apply_dense_kernel<Real, 4>), grid, dim3(256), 0, stream, psi, groups, tg, m); break;
        case 5: hipLaunchKernelGGL((apply_dense_kernel<Real, 5>), grid, dim3(256), 0, stream, psi, groups, tg, m); break;
        default: set_error("dense apply supports 1..5 targets, got %d", k); return E_UNSUPPORTED;
        }
        MIPT_HIP_CHECK(hipGetLastError());
        return OK;
    }

    // One whole trajectory (Circuit1D::try_cusv, cpp/include/mipt/circuit.hpp:240-282)
    // from |0...0>: every layer's blocked matrices and target tables are built
    // on the host first and staged to the device in one copy, so the block
    // applies run back to back with no host round trip; each measured layer
    // keeps the reference's structure (engine.hpp:293-368): FP64 histogram,
    // copy to the host, draw, collapse.
    int run_trajectory(int nq, int layers, int closed, int gx, const double *gates, const std::uint8_t *flags,
                       const double *uniforms, int block_targets, int mode, int chunk_bits, int *outcomes)
    {
        MIPT_HIP_TRY(mipt_hip_engine_reset_impl(nq));
        struct Blk
        {
            int k;
            std::size_t mat_off;  // elements (complex) into the staged matrices
        };
        std::vector<Blk> blocks;
        std::vector<int> layer_first(layers + 1, 0);
        std::vector<DenseTargets> tgs;
        std::vector<double> mats;  // interleaved complex, FP64 (narrowed below for complex64)
        std::size_t g = 0;
        std::vector<int> q0, q1;
        for (int t = 0; t < layers; ++t)
        {
            layer_first[t] = static_cast<int>(blocks.size());
            const int nb = geo_bond_count(nq, gx, t, closed);
            q0.resize(nb);
            q1.resize(nb);
            for (int b = 0; b < nb; ++b) geo_bond(nq, gx, t, b, closed, q0[b], q1[b]);
            MIPT_HIP_TRY(block_bonds(q0.data(), q1.data(), gates + 32 * g, nb, block_targets,
                                     [&](const int *tt, int k, const double *m) {
                                         DenseTargets tg{};
                                         tg.k = k;
                                         for (int j = 0; j < k; ++j) tg.sorted[j] = tt[j];
                                         std::sort(tg.sorted, tg.sorted + k);
                                         for (int a = 0; a < (1 << k); ++a)
                                             for (int j = 0; j < k; ++j)
                                                 if ((a >> j) & 1) tg.off[a] |= std::uint64_t{1} << tt[j];
                                         tgs.push_back(tg);
                                         blocks.push_back({k, mats.size() / 2});
                                         mats.insert(mats.end(), m, m + 2 * (std::size_t{1} << (2 * k)));
                                         return OK;
                                     }));
            g += nb;
        }
        layer_first[layers] = static_cast<int>(blocks.size());
        const std::size_t celem = fp64 ? 16 : 8;
        const std::size_t ncx = mats.size() / 2;
        if (ncx * celem > traj_mat_cap)
        {
            if (traj_mat) MIPT_HIP_CHECK(hipFree(traj_mat));
            traj_mat = nullptr;
            MIPT_HIP_CHECK(hipMalloc(&traj_mat, ncx * celem));
            traj_mat_cap = ncx * celem;
        }
        if (tgs.size() > traj_tg_cap)
        {
            if (traj_tg) MIPT_HIP_CHECK(hipFree(traj_tg));
            traj_tg = nullptr;
            MIPT_HIP_CHECK(hipMalloc(&traj_tg, tgs.size() * sizeof(DenseTargets)));
            traj_tg_cap = tgs.size();
        }
        std::vector<float> narrow;
        if (!fp64)
        {
            narrow.resize(mats.size());
            for (std::size_t i = 0; i < mats.size(); ++i) narrow[i] = static_cast<float>(mats[i]);
        }
        if (ncx)
            MIPT_HIP_CHECK(hipMemcpyAsync(traj_mat, fp64 ? static_cast<const void *>(mats.data()) : narrow.data(),
                                          ncx * celem, hipMemcpyHostToDevice, stream));
        if (!tgs.empty())
            MIPT_HIP_CHECK(hipMemcpyAsync(traj_tg, tgs.data(), tgs.size() * sizeof(DenseTargets), hipMemcpyHostToDevice,
                                          stream));
        std::vector<int> sites;
        std::vector<double> u;
        for (int t = 0; t < layers; ++t)
        {
            for (int bi = layer_first[t]; bi < layer_first[t + 1]; ++bi)
            {
                const Blk &b = blocks[bi];
                const void *m = static_cast<const char *>(traj_mat) + b.mat_off * celem;
                MIPT_HIP_TRY(fp64 ? launch_dense_at<double>(b.k, dim() >> b.k, traj_tg + bi, m)
                                  : launch_dense_at<float>(b.k, dim() >> b.k, traj_tg + bi, m));
            }
            sites.clear();
            u.clear();
            for (int q = 0; q < nq; ++q)
            {
                outcomes[static_cast<std::size_t>(t) * nq + q] = -1;
                if (flags[static_cast<std::size_t>(t) * nq + q])
                {
                    sites.push_back(q);
                    if (mode == 0) u.push_back(uniforms[static_cast<std::size_t>(t) * nq + q]);
                }
            }
            if (sites.empty()) continue;
            if (mode == 1) u.assign(uniforms + static_cast<std::size_t>(t) * nq, uniforms + static_cast<std::size_t>(t + 1) * nq);
            std::vector<int> o(sites.size());
            MIPT_HIP_TRY(fp64 ? measure_layer_t<double>(sites.data(), static_cast<int>(sites.size()), u.data(), mode,
                                                        chunk_bits, o.data(), nullptr, nullptr)
                              : measure_layer_t<float>(sites.data(), static_cast<int>(sites.size()), u.data(), mode,
                                                       chunk_bits, o.data(), nullptr, nullptr));
            for (std::size_t j = 0; j < sites.size(); ++j) outcomes[static_cast<std::size_t>(t) * nq + sites[j]] = o[j];
        }
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        return OK;
    }

    int mipt_hip_engine_reset_impl(int nn);

    template <typename Real>
    int measure_layer_t(const int *sites, int count, const double *uniforms, int mode, int chunk_bits, int *outcomes,
                        double *p1, double *surprisal)
    {
        double surpr = 0.0;
        std::uint64_t cmask = 0, cval = 0;
        double weight = 0.0, total0 = -1.0;
        int chunk = 0;
        for (int begin = 0; begin < count; begin += chunk_bits, ++chunk)
        {
            const int end = std::min(count, begin + chunk_bits);
            const int m = end - begin;
            std::uint64_t mmask = 0;
            for (int j = begin; j < end; ++j) mmask |= std::uint64_t{1} << sites[j];
            MIPT_HIP_TRY(conditional_hist<Real>(state, n, mmask, cmask, cval, stream, hs, hist));
            double total = 0.0;
            for (double v : hist) total += v;
            if (total0 < 0) total0 = total;
            if (!(total > 0.0) || !std::isfinite(total)) { set_error("Measurement layer found a zero-norm state vector."); return E_NORM; }
            std::uint32_t chosen = 0;
            if (mode == 0)
            {
                int st = 0;
                double pp[64];
                chosen = sequential_from_hist(hist.data(), m, uniforms + begin, pp, &st);
                if (st) { set_error("Measurement layer found a zero-weight branch."); return st; }
                if (p1) for (int j = 0; j < m; ++j) p1[begin + j] = pp[j];
            }
            else
            {
                const double draw = uniforms[chunk] * total;
                double running = 0.0;
                chosen = static_cast<std::uint32_t>(hist.size() - 1);
                for (std::size_t b = 0; b < hist.size(); ++b)
                {
                    running += hist[b];
                    if (draw < running) { chosen = static_cast<std::uint32_t>(b); break; }
                }
                if (!(hist[chosen] > 0.0))
                {
                    chosen = static_cast<std::uint32_t>(std::max_element(hist.begin(), hist.end()) - hist.begin());
                    if (!(hist[chosen] > 0.0)) { set_error("Measurement layer has no outcome with positive weight."); return E_PROB; }
                }
            }
            weight = hist[chosen];
            surpr += -std::log(weight / total);
            for (int j = 0; j < m; ++j)
            {
                const int bit = (chosen >> j) & 1u;
                outcomes[begin + j] = bit;
                cmask |= std::uint64_t{1} << sites[begin + j];
                if (bit) cval |= std::uint64_t{1} << sites[begin + j];
            }
        }
        if (count > 0)
        {
            if (!(weight > 0.0)) { set_error("Measurement selected a zero-weight branch."); return E_BRANCH; }
            MIPT_HIP_TRY(collapse<Real>(state, n, cmask, cval, 1.0 / std::sqrt(weight), stream));
        }
        if (surprisal) *surprisal = surpr;
        return OK;
    }
};

int PlainEngine::mipt_hip_engine_reset_impl(int nn)
{
    if (nn < 1 || nn > 34) { set_error("engine_reset: bad arguments"); return E_ARG; }
    if (!owned || n != nn)
    {
        if (owned) MIPT_HIP_CHECK(hipFree(owned));
        owned = nullptr;
        MIPT_HIP_CHECK(hipMalloc(&owned, (std::size_t{1} << nn) * (fp64 ? 16 : 8)));
    }
    state = owned;
    n = nn;
    const std::uint64_t d = std::uint64_t{1} << nn;
    if (fp64)
        hipLaunchKernelGGL(set_basis_kernel<double>, dim3(grid_for(d, 256)), dim3(256), 0, stream,
                           static_cast<double2 *>(state), d, std::uint64_t{0});
    else
        hipLaunchKernelGGL(set_basis_kernel<float>, dim3(grid_for(d, 256)), dim3(256), 0, stream,
                           static_cast<float2 *>(state), d, std::uint64_t{0});
    MIPT_HIP_CHECK(hipGetLastError());
    return OK;
}

// Plain device copy used to calibrate hardware traffic counters
// (FETCH_SIZE / WRITE_SIZE) against a known byte count.
__global__ void stream_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, std::size_t n)
{
    const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
    for (std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

} // namespace mipt_hip

using namespace mipt_hip;

struct mipt_hip_engine : PlainEngine
{
};

extern "C" {

const char *mipt_hip_last_error(void) { return mipt_hip::last_error(); }

int mipt_hip_stream_copy(void *dst, const void *src, uint64_t bytes, void *stream)
{
    if (!dst || !src || (bytes % 16) != 0) { set_error("stream_copy: bad arguments"); return E_ARG; }
    const std::size_t n = bytes / 16;
    hipLaunchKernelGGL(stream_copy_kernel, dim3(grid_for(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const float4 *>(src), static_cast<float4 *>(dst), n);
    MIPT_HIP_CHECK(hipGetLastError());
    MIPT_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return OK;
}
const char *mipt_hip_version(void) { return "mipt_hip 0.1 gfx950"; }

int mipt_hip_measure_collapse_f32(void *s, int n, int q, double u, int *o, double *p)
{
    return measure_collapse_impl<float>(s, n, q, u, o, p);
}
int mipt_hip_measure_collapse_f64(void *s, int n, int q, double u, int *o, double *p)
{
    return measure_collapse_impl<double>(s, n, q, u, o, p);
}
const char *mipt_hip_measure_collapse_last_error(void) { return t_measure_error; }
int mipt_hip_measure_collapse_f32_noalloc(void *s, int n, int q, double u, int *o, double *p)
{
    return measure_collapse_noalloc_impl<float>(s, n, q, u, o, p);
}

int mipt_hip_engine_create(int use_fp64, void *stream, mipt_hip_engine **out)
{
    if (!out) { set_error("engine_create: null out"); return E_ARG; }
    auto *e = new (std::nothrow) mipt_hip_engine();
    if (!e) return E_NOMEM;
    e->fp64 = use_fp64 != 0;
    e->stream = static_cast<hipStream_t>(stream);
    *out = e;
    return OK;
}

int mipt_hip_engine_destroy(mipt_hip_engine *e)
{
    delete e;
    return OK;
}

int mipt_hip_engine_adopt(mipt_hip_engine *e, void *state, int n)
{
    if (!e || !state || n < 1 || n > 34) { set_error("engine_adopt: bad arguments"); return E_ARG; }
    if (e->owned) { MIPT_HIP_CHECK(hipFree(e->owned)); e->owned = nullptr; }
    e->state = state;
    e->n = n;
    return OK;
}

int mipt_hip_engine_reset(mipt_hip_engine *e, int n)
{
    if (!e) { set_error("engine_reset: null engine"); return E_ARG; }
    return e->mipt_hip_engine_reset_impl(n);
}

int mipt_hip_engine_run_trajectory(mipt_hip_engine *e, int n, int layers, int closed, int grid_x, const double *gates,
                                   const uint8_t *flags, const double *uniforms, int block_targets, int mode,
                                   int chunk_bits, int *outcomes)
{
    if (!e || n < 2 || n > 34 || layers < 0 || !flags || !uniforms || !outcomes || (!gates && layers > 0) ||
        block_targets < 2 || block_targets > 5 || (mode != 0 && mode != 1) || chunk_bits < 1 || chunk_bits > 16 ||
        grid_x < 0 || (grid_x > 0 && (n % grid_x != 0 || grid_x < 2)))
    {
        set_error("engine_run_trajectory: bad arguments");
        return E_ARG;
    }
    for (std::size_t i = 0; i < static_cast<std::size_t>(layers) * n; ++i)
        if (flags[i] && mode == 0 && !(uniforms[i] >= 0.0 && uniforms[i] < 1.0))
        {
            set_error("engine_run_trajectory: uniform out of [0,1)");
            return E_ARG;
        }
    return e->run_trajectory(n, layers, closed ? 1 : 0, grid_x, gates, flags, uniforms, block_targets, mode, chunk_bits,
                             outcomes);
}

void *mipt_hip_engine_data(mipt_hip_engine *e) { return e ? e->state : nullptr; }

int mipt_hip_engine_apply_matrix(mipt_hip_engine *e, const int *targets, int k, const double *m)
{
    if (!e || !targets || !m) { set_error("apply_matrix: null argument"); return E_ARG; }
    return e->apply_matrix(targets, k, m);
}

int mipt_hip_engine_apply_bonds(mipt_hip_engine *e, const int *q0, const int *q1, const double *g, int count,
                                int block_targets)
{
    if (!e || !e->state || count < 0 || (count && (!q0 || !q1 || !g)) || block_targets < 2 || block_targets > 5)
    {
        set_error("apply_bonds: bad arguments");
        return E_ARG;
    }
    return e->apply_bonds(q0, q1, g, count, block_targets);
}

int mipt_hip_engine_measure_layer(mipt_hip_engine *e, const int *sites, int count, const double *u, int mode,
                                  int chunk_bits, int *outcomes, double *p1, double *surprisal)
{
    if (!e || !e->state || count < 0 || (count && (!sites || !u || !outcomes)) || (mode != 0 && mode != 1) ||
        chunk_bits < 1 || chunk_bits > 16 || (mode == 0 && chunk_bits > 16))
    {
        set_error("measure_layer: bad arguments");
        return E_ARG;
    }
    for (int j = 0; j < count; ++j)
    {
        if (sites[j] < 0 || sites[j] >= e->n || (j && sites[j] <= sites[j - 1])) { set_error("measure_layer: sites must be ascending and in range"); return E_ARG; }
        if (mode == 0 && !(u[j] >= 0.0 && u[j] < 1.0)) { set_error("measure_layer: uniform out of [0,1)"); return E_ARG; }
    }
    if (surprisal) *surprisal = 0.0;
    if (count == 0) return OK;
    return e->fp64 ? e->measure_layer_t<double>(sites, count, u, mode, chunk_bits, outcomes, p1, surprisal)
                   : e->measure_layer_t<float>(sites, count, u, mode, chunk_bits, outcomes, p1, surprisal);
}

int mipt_hip_engine_to_host(mipt_hip_engine *e, void *host)
{
    if (!e || !e->state || !host) { set_error("to_host: bad arguments"); return E_ARG; }
    MIPT_HIP_CHECK(hipMemcpyAsync(host, e->state, e->dim() * e->elem(), hipMemcpyDeviceToHost, e->stream));
    MIPT_HIP_CHECK(hipStreamSynchronize(e->stream));
    return OK;
}

int mipt_hip_engine_set_from_host(mipt_hip_engine *e, const void *host)
{
    if (!e || !e->state || !host) { set_error("set_from_host: bad arguments"); return E_ARG; }
    MIPT_HIP_CHECK(hipMemcpyAsync(e->state, host, e->dim() * e->elem(), hipMemcpyHostToDevice, e->stream));
    MIPT_HIP_CHECK(hipStreamSynchronize(e->stream));
    return OK;
}

int mipt_hip_engine_squared_norm(mipt_hip_engine *e, double *out)
{
    if (!e || !e->state || !out) { set_error("squared_norm: bad arguments"); return E_ARG; }
    std::vector<double> h;
    MIPT_HIP_TRY(e->fp64 ? conditional_hist<double>(e->state, e->n, 0, 0, 0, e->stream, e->hs, h)
                         : conditional_hist<float>(e->state, e->n, 0, 0, 0, e->stream, e->hs, h));
    *out = h[0];
    return OK;
}

} // extern "C"
