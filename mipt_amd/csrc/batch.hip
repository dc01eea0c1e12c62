// Batched MIPT trajectory engine for gfx950.
//
// Reference behaviour (paths relative to the reference repository root):
//   cpp/include/mipt/circuit.hpp:240-282   per-trajectory layer loop
//   cpp/include/mipt/circuits/haar.hpp:39-85, 226-305  Haar bonds, layers, flags
//   cpp/include/mipt/free_energy_measure.hpp:19-95     sequential Born sampling
//   cpp/include/mipt/cusv/engine.hpp:293-368           joint-chunk = sequential (chain rule)
//   cpp/include/mipt/probed.hpp:25-33                  probe Bell pair h(site); cx(site, n)
//
// Design (MI355X-first, not a translation):
//   * every state lives in HBM for the whole run (complex64, qubit q = bit q);
//   * chains whose state fits one workgroup's LDS (<= 14 qubits) run all their
//     layers in ONE launch, one workgroup per trajectory;
//   * longer chains are swept in tiles of 2^k amplitudes (k = 6 + RB + wave
//     bits): each tile keeps the low c physical bits plus a window of high bits
//     local, stages its amplitudes through LDS so that up to RB/2 bonds at a time
//     sit in registers, and applies the brickwork layer's bonds there.  A layer
//     takes as many sweeps as windows are needed to cover its bonds (2 for L=20);
//   * Born probabilities of the layer's measured sites are accumulated in FP64
//     in the last sweep (per-tile partials, fixed order), a finalize kernel
//     turns them into sequential outcomes, and the collapse with FP64
//     renormalisation is folded into the next layer's first sweep;
//   * trajectories are processed in groups of at most 4 GiB of state, two
//     groups in flight on two streams: a sweep launch over one group is then
//     up to 65536 workgroups, whose last partial round of four workgroups per
//     CU costs little (measured on MI355X at L=20 with matrix-core bonds:
//     448 MiB groups 179k layers/s, 896 MiB 188k, 1.75 GiB 192.2k, 3.5 GiB
//     192.6k; later, with layer fusion, 3.5 GiB 208.7k, 4 GiB 210.3k (4096
//     trajectories in eight even groups), 8 GiB 209.5k).
#include "batch.hpp"
#include "state_ops.hpp"

#include <algorithm>
#include <atomic>
#include <type_traits>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <map>
#include <numeric>
#include <vector>

namespace mipt_hip
{

// ============================================================== generation
// Gates of every (trajectory, layer, bond): Haar U(4) from the counter-based
// stream, narrowed to complex64 row-major.
__global__ void gen_gates_kernel(const TrajDesc *__restrict__ desc, const int *__restrict__ gen_list, int count,
                                 const std::uint32_t *__restrict__ bond_prefix_layers, float2 *__restrict__ gates,
                                 std::uint64_t total_items, const std::uint64_t *__restrict__ item_prefix)
{
    const std::uint64_t id = blockIdx.x * static_cast<std::uint64_t>(blockDim.x) + threadIdx.x;
    if (id >= total_items) return;
    // binary search the trajectory owning this item
    int lo = 0, hi = count - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) / 2;
        if (item_prefix[mid] <= id) lo = mid;
        else hi = mid - 1;
    }
    const TrajDesc d = desc[gen_list[lo]];
    const std::uint64_t local = id - item_prefix[lo];  // gate index within the trajectory
    // locate (layer, bond)
    const int P = period_layers(d.gx);
    const std::uint64_t per_period = geo_gate_prefix(d.L, d.gx, d.closed, P);
    const std::uint64_t period = local / per_period;
    std::uint64_t rem = local % per_period;
    int layer = static_cast<int>(P * period);
    for (int j = 0; j < P; ++j)
    {
        const std::uint64_t c = static_cast<std::uint64_t>(geo_bond_count(d.L, d.gx, j, d.closed));
        if (rem < c) break;
        rem -= c;
        ++layer;
    }
    double qr[16], qi[16];
    haar_u4_philox(d.seed, static_cast<std::uint32_t>(layer), static_cast<std::uint32_t>(rem), qr, qi);
    float2 *g = gates + (d.gate_off + local) * 16;
    for (int e = 0; e < 16; ++e) g[e] = make_float2(static_cast<float>(qr[e]), static_cast<float>(qi[e]));
    (void)bond_prefix_layers;
}

// Flags and measurement uniforms of every (trajectory, layer).
__global__ void gen_flags_kernel(const TrajDesc *__restrict__ desc, int count, std::uint32_t *__restrict__ flags,
                                 double *__restrict__ uniforms, std::uint64_t total_items,
                                 const std::uint64_t *__restrict__ layer_prefix)
{
    const std::uint64_t id = blockIdx.x * static_cast<std::uint64_t>(blockDim.x) + threadIdx.x;
    if (id >= total_items) return;
    int lo = 0, hi = count - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) / 2;
        if (layer_prefix[mid] <= id) lo = mid;
        else hi = mid - 1;
    }
    const TrajDesc d = desc[lo];
    const int t = static_cast<int>(id - layer_prefix[lo]);
    std::uint32_t mask = 0;
    for (int s = 0; s < d.L; ++s)
    {
        double a, b, m, unused;
        philox_uniform_pair(d.seed, static_cast<std::uint32_t>(t), static_cast<std::uint32_t>(s), 0, STREAM_FLAG, a, b);
        const double p = (t < d.p_zero_layers) ? 0.0 : d.p;
        const bool f = (p <= 0.0) ? false : (p >= 1.0 ? true : (a < p));
        philox_uniform_pair(d.seed, static_cast<std::uint32_t>(t), static_cast<std::uint32_t>(s), 0, STREAM_MEAS, m, unused);
        uniforms[d.unif_off + static_cast<std::size_t>(t) * d.L + s] = f ? m : 0.0;
        if (f) mask |= 1u << s;
    }
    flags[d.layer_off + t] = mask;
}

// ============================================================== small chains
// One workgroup per trajectory, the whole state in LDS, all requested layers
// in one launch.  Bonds one at a time (thread per group of 4 amplitudes),
// measured sites sequentially with fixed-order FP64 block reductions.
constexpr int SMALL_THREADS = 256;

__device__ inline double block_sum(double v, double *red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = SMALL_THREADS / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

constexpr std::uint32_t STREAM_ATCH = 0x41544348u;  // probe-attach reset uniforms

// Reference attach at the start of layer p_zero_layers (> 0), as
// cpp/include/mipt/probed/kernels.hpp:55-77: for j = 0, 1, ... in order,
// reset(site_j), h(site_j), cx(site_j, L + j), reference j being |0> until
// then.  The reset is a Born draw on the site (outcome 1 iff u < P1, u from
// the counter stream at (layer, site)), renormalised by 1/sqrt(branch weight)
// and flipped to 0; FP64 weights in a fixed reduction order.  256 threads,
// any buffer.
__device__ void attach_one(float2 *v, std::uint32_t dim, const TrajDesc &d, int lo, int hi, int layer, double *red,
                           int &status)
{
    const std::uint32_t S = 1u << lo, P = 1u << hi;
    double w = 0.0, w1 = 0.0;
    for (std::uint32_t i = threadIdx.x; i < dim; i += SMALL_THREADS)
    {
        const float2 x = v[i];
        const double m = static_cast<double>(x.x) * x.x + static_cast<double>(x.y) * x.y;
        w += m;
        if (i & S) w1 += m;
    }
    w = block_sum(w, red);
    w1 = block_sum(w1, red);
    double p = (w > 0.0) ? w1 / w : 0.0;
    p = p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);
    double u, unused;
    philox_uniform_pair(d.seed, static_cast<std::uint32_t>(layer), static_cast<std::uint32_t>(lo), 0, STREAM_ATCH, u,
                        unused);
    const int o = u < p ? 1 : 0;
    const double branch = o ? w1 : w - w1;
    if (!(branch > 0.0)) status = E_NORM;
    const float scale = static_cast<float>(0.70710678118654752440 / sqrt(branch > 0.0 ? branch : 1.0));
    for (std::uint32_t grp = threadIdx.x; grp < (dim >> 2); grp += SMALL_THREADS)
    {
        std::uint32_t base = ((grp >> lo) << (lo + 1)) | (grp & (S - 1u));
        base = ((base >> hi) << (hi + 1)) | (base & (P - 1u));
        float2 x = v[base | (o ? S : 0u)];
        x.x *= scale;
        x.y *= scale;
        v[base] = x;
        v[base | S | P] = x;
        v[base | S] = make_float2(0.f, 0.f);
        v[base | P] = make_float2(0.f, 0.f);
    }
    __syncthreads();
}

// Reference j attaches at layer attach + j * stagger (0: Bell pair at t = 0).
__device__ inline int attach_layer_of(const TrajDesc &d, int j) { return d.attach + j * d.stagger; }

// Attaches, in order, the references whose layer is t.
__device__ void attach_probe(float2 *v, std::uint32_t dim, const TrajDesc &d, int t, double *red, int &status)
{
#pragma unroll
    for (int j = 0; j < 4; ++j)  // constant indices keep psite in registers
    {
        if (j >= d.probes) break;
        if (attach_layer_of(d, j) == t) attach_one(v, dim, d, d.psite[j], d.L + j, t, red, status);
    }
}

__device__ inline bool attaches_at(const TrajDesc &d, int t)
{
    bool any = false;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (j < d.probes && attach_layer_of(d, j) == t) any = true;
    return any;
}

__global__ void __launch_bounds__(SMALL_THREADS) attach_kernel(DevArena a, const int *__restrict__ traj_list, int t)
{
    __shared__ double red[SMALL_THREADS];
    const int traj = traj_list[blockIdx.x];
    const TrajDesc d = a.desc[traj];
    if (t == 0 || !attaches_at(d, t)) return;  // uniform over the block
    int status = 0;
    attach_probe(a.states + d.state_off, 1u << d.nq, d, t, red, status);
    if (threadIdx.x == 0 && status) a.status[traj] = status;
}

__device__ inline float2 initial_amp(const TrajDesc &d, std::uint64_t i)
{
    if (d.probes > 0 && d.attach == 0)
    {
        // Bell pairs (site_j, L + j) of the references attached at t = 0:
        // amplitude 2^(-pairs/2) where every pair's two bits agree and all
        // other bits (later references included) are 0
        std::uint64_t rest = i;
        int pairs = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j)  // fully unrolled: constant indices keep psite in registers
        {
            if (j >= d.probes) break;
            if (attach_layer_of(d, j) != 0) continue;
            const std::uint64_t s = std::uint64_t{1} << d.psite[j], r = std::uint64_t{1} << (d.L + j);
            if (((i & s) != 0) != ((i & r) != 0)) return make_float2(0.f, 0.f);
            rest &= ~(s | r);
            ++pairs;
        }
        if (rest) return make_float2(0.f, 0.f);
        const float h = pairs == 1 ? 0.70710678118654752440f
                        : pairs == 2 ? 0.5f
                        : pairs == 3 ? 0.35355339059327376220f
                                     : 0.25f;
        return make_float2(h, 0.f);
    }
    return i == 0 ? make_float2(1.f, 0.f) : make_float2(0.f, 0.f);
}

__device__ inline std::uint64_t gate_prefix(const TrajDesc &d, int t) { return geo_gate_prefix(d.L, d.gx, d.closed, t); }

__global__ void __launch_bounds__(SMALL_THREADS) small_evolve_kernel(
    DevArena a, const int *__restrict__ traj_list, int layer_begin, int layer_end, int init)
{
    extern __shared__ float2 st[];
    __shared__ double red[SMALL_THREADS];
    const int traj = traj_list[blockIdx.x];
    const TrajDesc d = a.desc[traj];
    const int nq = d.nq;
    const std::uint32_t dim = 1u << nq;
    float2 *psi = a.states + d.state_off;
    const bool gen = init && a.init[traj];
    for (std::uint32_t i = threadIdx.x; i < dim; i += SMALL_THREADS) st[i] = gen ? initial_amp(d, i) : psi[i];
    __syncthreads();
    double surpr = 0.0;
    int status = 0;
    for (int t = layer_begin; t < layer_end; ++t)
    {
        if (t > 0 && attaches_at(d, t)) attach_probe(st, dim, d, t, red, status);
        const int nb = geo_bond_count(d.L, d.gx, t, d.closed);
        const std::uint64_t gbase = d.gate_off + gate_prefix(d, t);
        for (int b = 0; b < nb; ++b)
        {
            int q0, q1;
            geo_bond(d.L, d.gx, t, b, d.closed, q0, q1);
            const float2 *g = a.gates + (gbase + b) * 16;
            float ur[16], ui[16];
            for (int e = 0; e < 16; ++e)
            {
                ur[e] = g[e].x;
                ui[e] = g[e].y;
            }
            const int lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
            for (std::uint32_t grp = threadIdx.x; grp < (dim >> 2); grp += SMALL_THREADS)
            {
                std::uint32_t base = ((grp >> lo) << (lo + 1)) | (grp & ((1u << lo) - 1));
                base = ((base >> hi) << (hi + 1)) | (base & ((1u << hi) - 1));
                const std::uint32_t i1 = base | (1u << q0), i2 = base | (1u << q1), i3 = i1 | i2;
                const float2 x0 = st[base], x1 = st[i1], x2 = st[i2], x3 = st[i3];
                float2 y[4];
#pragma unroll
                for (int r = 0; r < 4; ++r)
                {
                    float yr = ur[4 * r] * x0.x - ui[4 * r] * x0.y;
                    float yi = ur[4 * r] * x0.y + ui[4 * r] * x0.x;
                    yr = fmaf(ur[4 * r + 1], x1.x, fmaf(-ui[4 * r + 1], x1.y, yr));
                    yi = fmaf(ur[4 * r + 1], x1.y, fmaf(ui[4 * r + 1], x1.x, yi));
                    yr = fmaf(ur[4 * r + 2], x2.x, fmaf(-ui[4 * r + 2], x2.y, yr));
                    yi = fmaf(ur[4 * r + 2], x2.y, fmaf(ui[4 * r + 2], x2.x, yi));
                    yr = fmaf(ur[4 * r + 3], x3.x, fmaf(-ui[4 * r + 3], x3.y, yr));
                    yi = fmaf(ur[4 * r + 3], x3.y, fmaf(ui[4 * r + 3], x3.x, yi));
                    y[r] = make_float2(yr, yi);
                }
                st[base] = y[0];
                st[i1] = y[1];
                st[i2] = y[2];
                st[i3] = y[3];
            }
            __syncthreads();
        }
        const std::uint32_t fl = a.flags[d.layer_off + t];
        std::uint32_t outc = 0, flg = 0;
        for (int s = 0; s < d.L; ++s)
        {
            if (!((fl >> s) & 1u)) continue;
            double w = 0.0, w1 = 0.0;
            for (std::uint32_t i = threadIdx.x; i < dim; i += SMALL_THREADS)
            {
                const float2 v = st[i];
                const double x = static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y;
                w += x;
                if ((i >> s) & 1u) w1 += x;
            }
            w = block_sum(w, red);
            w1 = block_sum(w1, red);
            double p = (w > 0.0) ? w1 / w : 0.0;
            p = p < 0.0 ? 0.0 : (p > 1.0 ? 1.0 : p);
            const double u = a.uniforms[d.unif_off + static_cast<std::size_t>(t) * d.L + s];
            const int o = u < p ? 1 : 0;
            const double branch = o ? w1 : w - w1;
            if (!(w > 0.0) || !(branch > 0.0)) status = E_NORM;
            if (threadIdx.x == 0) a.p1[d.unif_off + static_cast<std::size_t>(t) * d.L + s] = p;
            if (o) outc |= 1u << s;
            if (fabs(u - p) < 1e-6) flg |= 1u << s;
            surpr += -log((o ? p : 1.0 - p));
            const float scale = static_cast<float>(1.0 / sqrt(branch > 0.0 ? branch : 1.0));
            for (std::uint32_t i = threadIdx.x; i < dim; i += SMALL_THREADS)
            {
                float2 v = st[i];
                if (((i >> s) & 1u) == static_cast<std::uint32_t>(o))
                {
                    v.x *= scale;
                    v.y *= scale;
                }
                else
                    v = make_float2(0.f, 0.f);
                st[i] = v;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0)
        {
            a.outcomes[d.layer_off + t] = outc;
            a.flagged[d.layer_off + t] = flg;
            a.coll[d.layer_off + t] = Collapse{0, 0, 1.0, 1.0};
        }
    }
    for (std::uint32_t i = threadIdx.x; i < dim; i += SMALL_THREADS) psi[i] = st[i];
    if (threadIdx.x == 0)
    {
        a.surprisal[traj] += surpr;
        if (status) a.status[traj] = status;
    }
}

// ============================================================== bulk readout
// Measurement records of every trajectory, int8 [count][nt][L]: -1 where the
// site was not measured, else the outcome.
__global__ void records_kernel(DevArena a, int t0, int nt, int L, std::size_t total, std::int8_t *__restrict__ out)
{
    const std::size_t id = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
    if (id >= total) return;
    const std::size_t per = static_cast<std::size_t>(nt) * L;
    const int traj = static_cast<int>(id / per);
    const int rem = static_cast<int>(id - static_cast<std::size_t>(traj) * per);
    const int t = t0 + rem / L, s = rem % L;
    const TrajDesc d = a.desc[traj];
    const std::uint32_t f = a.flags[d.layer_off + t], o = a.outcomes[d.layer_off + t];
    out[id] = ((f >> s) & 1u) ? static_cast<std::int8_t>((o >> s) & 1u) : static_cast<std::int8_t>(-1);
}

// Probe Bloch vectors [count][3] ((2 Re rho01, -2 Im rho01, rho00 - rho11) / tr,
// rho01 = sum psi[..0..] conj(psi[..1..]) over the probe qubit L; NaN without a
// probe), FP64 sums in a fixed order, one block per trajectory.
__global__ void __launch_bounds__(SMALL_THREADS) probe_bloch_kernel(DevArena a, double *__restrict__ out)
{
    __shared__ double red[SMALL_THREADS];
    const int traj = blockIdx.x;
    const TrajDesc d = a.desc[traj];
    if (d.probe_site < 0)
    {
        if (threadIdx.x < 3) out[traj * 3 + threadIdx.x] = NAN;
        return;
    }
    const float2 *psi = a.states + d.state_off;
    const std::uint32_t half = 1u << (d.nq - 1), P = 1u << d.L;
    double a00 = 0, a11 = 0, re = 0, im = 0;
    for (std::uint32_t e = threadIdx.x; e < half; e += SMALL_THREADS)
    {
        const std::uint32_t i0 = (e & (P - 1u)) | ((e & ~(P - 1u)) << 1);
        const float2 x = psi[i0], y = psi[i0 | P];
        a00 += static_cast<double>(x.x) * x.x + static_cast<double>(x.y) * x.y;
        a11 += static_cast<double>(y.x) * y.x + static_cast<double>(y.y) * y.y;
        re += static_cast<double>(x.x) * y.x + static_cast<double>(x.y) * y.y;
        im += static_cast<double>(x.y) * y.x - static_cast<double>(x.x) * y.y;
    }
    a00 = block_sum(a00, red);
    a11 = block_sum(a11, red);
    re = block_sum(re, red);
    im = block_sum(im, red);
    if (threadIdx.x == 0)
    {
        const double tr = a00 + a11;
        out[traj * 3 + 0] = 2.0 * re / tr;
        out[traj * 3 + 1] = -2.0 * im / tr;
        out[traj * 3 + 2] = (a00 - a11) / tr;
    }
}

// Joint density matrix of a trajectory's k references (qubits L..L+k-1,
// reduced-basis bit j = reference j): rho[r][c] = sum over the chain's 2^L
// basis states e of psi[e + r 2^L] conj(psi[e + c 2^L]).  The lower
// triangle is cut into T x T tiles (T = min(2^k, 4)); block (tile, traj)
// loads its T (diagonal tile) or 2T rows once per e and forms the tile's
// products from registers; FP64 sums over a fixed thread order.
// out[traj][q][re, im] with q = r(r+1)/2 + c for r >= c, not normalised.
__global__ void __launch_bounds__(SMALL_THREADS) probe_rdm_kernel(DevArena a, int D, double *__restrict__ out)
{
    __shared__ double red[SMALL_THREADS];
    const int traj = blockIdx.y;
    const TrajDesc d = a.desc[traj];
    const int T = D < 4 ? D : 4;
    int tq = blockIdx.x, I = 0;
    while (tq > I)
    {
        ++I;
        tq -= I;
    }
    const int J = tq;
    const float2 *psi = a.states + d.state_off;
    const std::uint32_t n = 1u << d.L;
    double re[4][4], im[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) re[r][c] = im[r][c] = 0.0;
    for (std::uint32_t e = threadIdx.x; e < n; e += SMALL_THREADS)
    {
        float2 x[4], y[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
        {
            x[t] = t < T ? psi[e + (static_cast<std::uint32_t>(I * T + t) << d.L)] : make_float2(0.f, 0.f);
            y[t] = (t < T && I != J) ? psi[e + (static_cast<std::uint32_t>(J * T + t) << d.L)] : x[t];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
                re[r][c] += static_cast<double>(x[r].x) * y[c].x + static_cast<double>(x[r].y) * y[c].y;
                im[r][c] += static_cast<double>(x[r].y) * y[c].x - static_cast<double>(x[r].x) * y[c].y;
            }
    }
    const int npair = D * (D + 1) / 2;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
            const int R = I * T + r, C = J * T + c;
            if (r >= T || c >= T || C > R) continue;  // uniform over the block
            const double sr = block_sum(re[r][c], red);
            const double si = block_sum(im[r][c], red);
            if (threadIdx.x == 0)
            {
                const std::size_t q = static_cast<std::size_t>(traj) * npair + R * (R + 1) / 2 + C;
                out[2 * q] = sr;
                out[2 * q + 1] = si;
            }
        }
}

// Three-site reduced density matrices (8 x 8) of listed subsystems, one block
// per (trajectory, subsystem): rho[r][c] = sum over the environment of
// psi[e|r] conj(psi[e|c]), reduced-basis bit k = subsystem qubit k, FP64
// sums of the complex64 amplitudes in a fixed order, then divided by the
// trace.  Output [traj][subsystem][8][8][re, im] (the record payload of
// cpp/include/mipt/density.hpp:40-48).
__global__ void __launch_bounds__(SMALL_THREADS) rho3_kernel(DevArena a, int first, int nsub,
                                                            const int *__restrict__ subs, double *__restrict__ out)
{
    __shared__ double acc_s[72];
    const int traj = first + blockIdx.x / nsub, sub = blockIdx.x % nsub;
    const TrajDesc d = a.desc[traj];
    const float2 *psi = a.states + d.state_off;
    const int q[3] = {subs[3 * sub], subs[3 * sub + 1], subs[3 * sub + 2]};
    int srt[3] = {q[0], q[1], q[2]};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2 - i; ++j)
            if (srt[j] > srt[j + 1])
            {
                const int t = srt[j];
                srt[j] = srt[j + 1];
                srt[j + 1] = t;
            }
    std::uint32_t off[8];
    for (int k = 0; k < 8; ++k)
        off[k] = ((k & 1) ? (1u << q[0]) : 0u) | ((k & 2) ? (1u << q[1]) : 0u) | ((k & 4) ? (1u << q[2]) : 0u);
    // lower triangle (36 complex entries) per thread
    double re[36], im[36];
#pragma unroll
    for (int i = 0; i < 36; ++i) re[i] = im[i] = 0.0;
    const std::uint32_t nenv = 1u << (d.nq - 3);
    for (std::uint32_t e = threadIdx.x; e < nenv; e += SMALL_THREADS)
    {
        std::uint32_t base = e;
        for (int j = 0; j < 3; ++j)
            base = ((base >> srt[j]) << (srt[j] + 1)) | (base & ((1u << srt[j]) - 1u));
        double xr[8], xi[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            const float2 v = psi[base | off[k]];
            xr[k] = v.x;
            xi[k] = v.y;
        }
        int t = 0;
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int c = 0; c <= r; ++c, ++t)
            {
                re[t] += xr[r] * xr[c] + xi[r] * xi[c];
                im[t] += xi[r] * xr[c] - xr[r] * xi[c];
            }
    }
    // wave sums by shuffles (fixed tree), then the four waves in order
    __shared__ double wsum[SMALL_THREADS / 64][72];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < 36; ++i)
    {
        double vr = re[i], vi = im[i];
#pragma unroll
        for (int sh = 32; sh > 0; sh >>= 1)
        {
            vr += __shfl_down(vr, sh, 64);
            vi += __shfl_down(vi, sh, 64);
        }
        if (lane == 0)
        {
            wsum[wave][2 * i] = vr;
            wsum[wave][2 * i + 1] = vi;
        }
    }
    __syncthreads();
    if (threadIdx.x < 72)
    {
        double t = 0.0;
        for (int w = 0; w < SMALL_THREADS / 64; ++w) t += wsum[w][threadIdx.x];
        acc_s[threadIdx.x] = t;
    }
    __syncthreads();
    if (threadIdx.x < 64)
    {
        const int r = threadIdx.x >> 3, c = threadIdx.x & 7;
        double tr = 0.0;
        for (int k = 0; k < 8; ++k) tr += acc_s[2 * (k * (k + 1) / 2 + k)];
        const int lo = r >= c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r;
        const double vr = acc_s[2 * lo], vi = (r >= c ? 1.0 : -1.0) * acc_s[2 * lo + 1];
        const double f = tr > 0.0 ? 1.0 / tr : 1.0;
        double *o = out + (static_cast<std::size_t>(blockIdx.x) * 64 + r * 8 + c) * 2;
        o[0] = vr * f;
        o[1] = vi * f;
    }
}

// ============================================================== tiled sweeps
// u = in-tile index (k bits).  u bits 0..c-1 <-> physical bits 0..c-1; u bits
// c..k-1 <-> physical bits s..s+k-c-1 (the window); the other physical bits
// enumerate the tiles.  A "mapping" assigns each u bit to a register bit (RB
// of them, so a thread holds 2^RB amplitudes), a lane bit (6) or a wave bit.
// M0 (global loads/stores and the histogram; m0_u): register bit 0 = u 0,
// lanes = u 1..6, register bits 1.. = u 7..5+RB, waves = the rest, so every
// 16-byte load/store instruction covers runs of 2^c consecutive amplitudes.
struct StagePlan
{
    // 32-bit fields: a uniform runtime index into the kernel arguments then
    // compiles to scalar loads (sub-dword fields force vector loads).
    int reg_u[6];
    int lane_u[6];
    int wave_u[4];
    int nb;        // bonds in this stage
    // bond number (within the layer), -1 none: bond[0] on register bits
    // (1,2), or with zpair 1 / 2 a bond of qubit 0 on bits (0,1) / (1,0)
    // (q0, q1); bond[1] on bits (3,4).  Register bit 0 always holds u bit 0
    // (qubit 0), on which the LDS exchanges split the tile in halves.
    int bond[2];
    int zpair;
    int use_lds;   // 0: M0 registers (no exchange)
};

struct PassPlan
{
    std::uint64_t rest_mask;  // physical bits enumerating tiles
    int ntiles;
    int k, c, s;
    int nstages;
    StagePlan st[6];
    int collapse;      // apply the previous layer's pending collapse at load
    int init;          // generate the initial state instead of loading
    int hist;          // accumulate the chunk histogram after the gates
    int store;         // write the tile back
    int chunk;         // histogram chunk index (sites chunk*MAX_HIST_BITS ...)
    int cond;          // restrict the histogram to the decided outcomes of the layer
    int zskip;         // leave tiles on the collapsed-out branch untouched (the next sweep zeroes them)
    std::uint64_t bonded;  // sites whose bonds this sweep applies
    std::uint64_t pmask;   // previous layer's collapsed sites to zero again at load (no rescale)
    // layer fusion: this is the layer's last sweep, launched over the
    // trajectories that fuse this layer into the next, which continue with
    // the next layer's first sweep (PassArgs::next, the same window) on the
    // tile
    int fuse;
    // with layer fusion a list entry ~i marks trajectory i as having run this
    // layer's first sweep fused into the previous layer's last, before that
    // layer's outcomes were drawn: with pend (the layer's second sweep) such
    // a trajectory gets the previous collapse applied at load, as a first
    // sweep applies it
    int pend;
    int pbase;  // the launch's i-th trajectory writes histogram partial record pbase + i
};

struct PassArgs
{
    DevArena a;
    const int *traj_list;  // group trajectories
    int ntraj;             // trajectories in the group (tiles = ntraj * plan.ntiles)
    int layer;
    PassPlan plan;
    PassPlan next;         // with plan.fuse: the next layer's first sweep (same window; its stages are used)
};

__device__ __forceinline__ std::uint32_t swz(std::uint32_t u)
{
    return u ^ (((u >> 4) ^ (u >> 8) ^ (u >> 12)) & 15u);
}

// LDS word of in-tile index u: the low nibble is XORed with the higher
// nibbles so that lanes walking any 6 u bits with distinct residues mod 4 hit
// distinct banks.  swz is GF(2)-linear: swz(a ^ b) = swz(a) ^ swz(b).
// 4x4 complex gate (row-major, index 2*b(q1)+b(q0)) on one register group
// (registers i0..i3 hold the basis states q1q0 = 00, 01, 10, 11) on the VALU,
// with the gate's coefficients uniform over the wave (scalar registers).
template <int NV, int i0, int i1, int i2, int i3>
__device__ __forceinline__ void group_valu(float2 (&v)[NV], const float (&ur)[16], const float (&ui)[16])
{
    const float2 x0 = v[i0], x1 = v[i1], x2 = v[i2], x3 = v[i3];
    float2 y[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
    {
        float yr = ur[4 * r] * x0.x;
        float yi = ur[4 * r] * x0.y;
        yr = fmaf(-ui[4 * r], x0.y, yr);
        yi = fmaf(ui[4 * r], x0.x, yi);
        yr = fmaf(ur[4 * r + 1], x1.x, yr);
        yi = fmaf(ur[4 * r + 1], x1.y, yi);
        yr = fmaf(-ui[4 * r + 1], x1.y, yr);
        yi = fmaf(ui[4 * r + 1], x1.x, yi);
        yr = fmaf(ur[4 * r + 2], x2.x, yr);
        yi = fmaf(ur[4 * r + 2], x2.y, yi);
        yr = fmaf(-ui[4 * r + 2], x2.y, yr);
        yi = fmaf(ui[4 * r + 2], x2.x, yi);
        yr = fmaf(ur[4 * r + 3], x3.x, yr);
        yi = fmaf(ur[4 * r + 3], x3.y, yi);
        yr = fmaf(-ui[4 * r + 3], x3.y, yr);
        yi = fmaf(ui[4 * r + 3], x3.x, yi);
        y[r] = make_float2(yr, yi);
    }
    v[i0] = y[0];
    v[i1] = y[1];
    v[i2] = y[2];
    v[i3] = y[3];
}

typedef float floatx4_t __attribute__((ext_vector_type(4)));

// The same group on the matrix cores.  v_mfma_f32_4x4x1_16b_f32 adds, in
// every block of four lanes 4b..4b+3, A(lane 4m + i) * B(lane l) to register
// i of lane l, where m = ABID is the block whose A values are broadcast to
// all 16 blocks (CBSZ = 4).  B is the lane's own amplitude component and the
// A column of block m = 8 T + 2 c + r is column (c, r) of a real 4 x 8 block
// T of the gate (r = 0: the coefficients of Re x_c, 1: of Im x_c), so the
// whole gate sits in ONE register kv (16 columns x 4 rows) and eight steps
// per block leave four outputs of the lane's own group in its own registers:
// no data moves between lanes and every product is used.  The rows of block
// T are (Re y_a, Im y_a, Re y_b, Im y_b) with (a, b) = (2T, 2T + 1), or (T,
// T + 2) with PQ1, so each accumulator holds two complex amplitudes in place;
// the pairing is chosen so that the two share register bit 1 where the bond
// allows, which lets the register allocator keep one layout of the tile
// across bond forms.
#define MIPT_MFMA_STEP(acc, xv, m) acc = __builtin_amdgcn_mfma_f32_4x4x1f32(kv, xv, acc, 4, m, 0)
template <int NV, int i0, int i1, int i2, int i3, bool PQ1>
__device__ __forceinline__ void group_mfma(float2 (&v)[NV], float kv)
{
    const float2 x0 = v[i0], x1 = v[i1], x2 = v[i2], x3 = v[i3];
    floatx4_t t0 = {0.f, 0.f, 0.f, 0.f}, t1 = {0.f, 0.f, 0.f, 0.f};
    MIPT_MFMA_STEP(t0, x0.x, 0);
    MIPT_MFMA_STEP(t1, x0.x, 8);
    MIPT_MFMA_STEP(t0, x0.y, 1);
    MIPT_MFMA_STEP(t1, x0.y, 9);
    MIPT_MFMA_STEP(t0, x1.x, 2);
    MIPT_MFMA_STEP(t1, x1.x, 10);
    MIPT_MFMA_STEP(t0, x1.y, 3);
    MIPT_MFMA_STEP(t1, x1.y, 11);
    MIPT_MFMA_STEP(t0, x2.x, 4);
    MIPT_MFMA_STEP(t1, x2.x, 12);
    MIPT_MFMA_STEP(t0, x2.y, 5);
    MIPT_MFMA_STEP(t1, x2.y, 13);
    MIPT_MFMA_STEP(t0, x3.x, 6);
    MIPT_MFMA_STEP(t1, x3.x, 14);
    MIPT_MFMA_STEP(t0, x3.y, 7);
    MIPT_MFMA_STEP(t1, x3.y, 15);
    // block T holds outputs (2T, 2T + 1), or with PQ1 (T, T + 2)
    constexpr int o1 = PQ1 ? i2 : i1, o2 = PQ1 ? i1 : i2;
    v[i0] = make_float2(t0[0], t0[1]);
    v[o1] = make_float2(t0[2], t0[3]);
    v[o2] = make_float2(t1[0], t1[1]);
    v[i3] = make_float2(t1[2], t1[3]);
}

#undef MIPT_MFMA_STEP

// Matrix-core coefficient table of a sweep in LDS: per bond slot (2 per
// stage: slot 2 si + j is bond[j] of stage si) 8 entries of 8 floats, entry
// 2 a + p for output a and part p (0: real, 1: imaginary): (U_r[a][c],
// -U_i[a][c]) for the real part, (U_i[a][c], U_r[a][c]) for the imaginary
// part, c = 0..3 -- the coefficients of (Re x_c, Im x_c).  256 bytes per slot.
constexpr int MX_SLOT_BYTES = 256;
constexpr int MX_MAX_SLOTS = 12;

// Writes the table of the sweep's bonds (gates gl) into tab: wave w fills
// slots w, w + NT/64, ...; lane e < 16 converts gate entry e = 4 a + c.
// The stage records are read with a wave-uniform slot index (scalar loads).
template <int NT>
__device__ __forceinline__ void fill_mx_table(const PassPlan &pl, const float2 *__restrict__ gl, float *tab)
{
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int e = static_cast<int>(threadIdx.x & 63u);
    for (int slot = w; slot < 2 * pl.nstages; slot += NT / 64)
    {
        const int b = pl.st[slot >> 1].bond[slot & 1];
        if (b >= 0 && e < 16)
        {
            const float2 g = gl[16 * b + e];
            const int a = e >> 2, c = e & 3;
            float *row = tab + slot * (MX_SLOT_BYTES / 4) + 2 * a * 8 + 2 * c;
            *reinterpret_cast<float2 *>(row) = make_float2(g.x, -g.y);
            *reinterpret_cast<float2 *>(row + 8) = make_float2(g.y, g.x);
        }
    }
}

// The groups of a bond on register bits RA, RC, enumerated at compile time
// (every register index a constant): group number GI starting at register
// BASE goes to the VALU while GI < VG, else to the matrix cores.
template <int NV, int RA, int RC, int VG, int BASE, int GI>
__device__ __forceinline__ void bond_groups(float2 (&v)[NV], const float (&ur)[16], const float (&ui)[16], float kv)
{
    if constexpr (BASE < NV)
    {
        if constexpr ((BASE & ((1 << RA) | (1 << RC))) != 0)
            bond_groups<NV, RA, RC, VG, BASE + 1, GI>(v, ur, ui, kv);
        else
        {
            constexpr int i1 = BASE | (1 << RA), i2 = BASE | (1 << RC), i3 = i1 | i2;
            if constexpr (GI < VG) group_valu<NV, BASE, i1, i2, i3>(v, ur, ui);
            else group_mfma<NV, BASE, i1, i2, i3, RC == 1>(v, kv);
            // a VALU-only bond keeps one group in flight at a time (its eight
            // FMA chains are enough to fill the pipe), which bounds registers
            if constexpr (VG == NV / 4) __builtin_amdgcn_sched_barrier(0);
            bond_groups<NV, RA, RC, VG, BASE + 1, GI + 1>(v, ur, ui, kv);
        }
    }
}

// A bond on register bits (RA <- q0, RC <- q1): its NV/4 groups are split
// between the VALU and the matrix cores by MX (0: all VALU; 1: NV/16 groups
// on the VALU, the rest on the matrix cores; 2: all on the matrix cores).
// A matrix-core group costs 16 MFMAs of 4x4x1 (256 products each, all
// useful), a VALU group 64 FMAs.  tab: the bond's slot in the coefficient
// table (matrix-core forms).
template <int NV, int RA, int RC, int MX>
__device__ __forceinline__ void bond_on_regs(float2 (&v)[NV], const float2 *__restrict__ g, const float *tab, int lane)
{
    constexpr int VG = MX == 0 ? NV / 4 : (MX == 1 ? NV / 16 : 0);
    float ur[16], ui[16];
    if constexpr (VG > 0)
    {
#pragma unroll
        for (int e = 0; e < 16; ++e)
        {
            const float2 x = g[e];
            ur[e] = x.x;
            ui[e] = x.y;
        }
    }
    float kv = 0.f;
    if constexpr (VG < NV / 4)
    {
        // lane 4 m + i holds row i of column m = 8 T + 2 c + r: table entry
        // 2 a + (i & 1), element 2 c + r, with a = 2 T + (i >> 1), or
        // T + 2 (i >> 1) when the blocks pair outputs across q1 (RC == 1)
        const int i = lane & 3, m = lane >> 2;
        const int T = m >> 3;
        const int a = RC == 1 ? T + 2 * (i >> 1) : 2 * T + (i >> 1);
        kv = tab[(2 * a + (i & 1)) * 8 + (m & 7)];
    }
    bond_groups<NV, RA, RC, VG, 0, 0>(v, ur, ui, kv);
}

// A stage's bonds: bond[0] on register bits (1,2) or, for a bond of qubit 0,
// (0,1) / (1,0) (zpair 1 / 2); bond[1] on (3,4).  tab: the stage's first
// table slot.  ZP: 0 -- the stage has no bond of qubit 0; 1 / 2 -- its
// bond[0] is the zpair 1 / 2 form; 3 -- any form, chosen at run time.  A
// sweep kernel that knows the form at compile time keeps one register
// layout of the tile across its stage loop (no join of differently laid out
// bond forms), which saves registers.
template <int NV, int MX, int ZP>
__device__ __forceinline__ void apply_stage_gates(float2 (&v)[NV], const StagePlan &sp, const float2 *gl,
                                                  const float *tab, int lane)
{
    static_assert(NV == 32 || NV == 64, "register tiles hold 32 or 64 amplitudes");
    if (sp.bond[0] >= 0)
    {
        if constexpr (ZP == 1) bond_on_regs<NV, 0, 1, MX>(v, gl + 16 * sp.bond[0], tab, lane);
        else if constexpr (ZP == 2) bond_on_regs<NV, 1, 0, MX>(v, gl + 16 * sp.bond[0], tab, lane);
        else if constexpr (ZP == 0) bond_on_regs<NV, 1, 2, MX>(v, gl + 16 * sp.bond[0], tab, lane);
        else
        {
            // The three forms differ only in which registers they touch; the
            // empty markers keep the optimiser from merging them into one body
            // that would address the tile through pointer selects (which moves
            // it to scratch).
            if (sp.zpair == 1)
            {
                bond_on_regs<NV, 0, 1, MX>(v, gl + 16 * sp.bond[0], tab, lane);
                asm volatile("; bond slot (0,1)");
            }
            else if (sp.zpair == 2)
            {
                bond_on_regs<NV, 1, 0, MX>(v, gl + 16 * sp.bond[0], tab, lane);
                asm volatile("; bond slot (1,0)");
            }
            else
            {
                bond_on_regs<NV, 1, 2, MX>(v, gl + 16 * sp.bond[0], tab, lane);
                asm volatile("; bond slot (1,2)");
            }
        }
    }
    if (sp.bond[1] >= 0) bond_on_regs<NV, 3, 4, MX>(v, gl + 16 * sp.bond[1], tab + MX_SLOT_BYTES / 4, lane);
}

// LDS addressing.  u bit 0 (qubit 0) is register bit 0 in every mapping, and
// an exchange moves the tile in two halves (u bit 0 = 0, then 1) through a
// buffer of 2^(k-1) words: u's word is swz(u >> 1), so the buffer is half
// the tile and more workgroups fit on a CU.  Address pieces of a mapping:
// thread part T3 = 8 swz((lane and wave bits) >> 1) and per-register-bit
// terms Q3[r] = 8 swz((1 << reg_u[r]) >> 1) for r >= 1 (Q3[0] unused).
template <int WB>
__device__ __forceinline__ std::uint32_t thread_term(const int *lane_u, const int *wave_u, int lane, int wave)
{
    std::uint32_t o = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b) o |= static_cast<std::uint32_t>((lane >> b) & 1) << lane_u[b];
#pragma unroll
    for (int b = 0; b < WB; ++b) o |= static_cast<std::uint32_t>((wave >> b) & 1) << wave_u[b];
    return swz(o >> 1);
}

// LDS byte address of register j: T3 ^ (XOR of Q3[r] over the set bits r >= 1 of j).
template <int NV>
__device__ __forceinline__ std::uint32_t reg_addr(std::uint32_t T3, const std::uint32_t (&Q3)[6], int j)
{
    std::uint32_t s = 0;
#pragma unroll
    for (int r = 1; r < 6; ++r)
        if ((j >> r) & 1) s ^= Q3[r];
    return T3 ^ s;
}

// One half of an exchange: registers j with j & 1 == HALF (u bit 0 = HALF).
template <int NV, int HALF>
__device__ __forceinline__ void lds_exchange_half(char *lds, float2 (&v)[NV], std::uint32_t T3s,
                                                  const std::uint32_t (&Q3s)[6], std::uint32_t T3d,
                                                  const std::uint32_t (&Q3d)[6], bool dirty)
{
    if (dirty) __syncthreads();
#pragma unroll
    for (int i = 0; i < NV / 2; ++i)
        *reinterpret_cast<float2 *>(lds + reg_addr<NV>(T3s, Q3s, 2 * i + HALF)) = v[2 * i + HALF];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV / 2; ++i)
        v[2 * i + HALF] = *reinterpret_cast<const float2 *>(lds + reg_addr<NV>(T3d, Q3d, 2 * i + HALF));
}

// Moves the tile from one mapping (T3s, Q3s) to another (T3d, Q3d) through
// LDS, half by half; dirty: an earlier read of the buffer may be pending.
template <int NV>
__device__ __forceinline__ void lds_exchange(char *lds, float2 (&v)[NV], std::uint32_t T3s, const std::uint32_t (&Q3s)[6],
                                             std::uint32_t T3d, const std::uint32_t (&Q3d)[6], bool &dirty)
{
    lds_exchange_half<NV, 0>(lds, v, T3s, Q3s, T3d, Q3d, dirty);
    lds_exchange_half<NV, 1>(lds, v, T3s, Q3s, T3d, Q3d, true);
    dirty = true;
}

// The M0 mapping of a tile (global loads and stores, the histogram): lane bit
// b <-> u bit b + 1, register bit 0 <-> u bit 0, register bit r >= 1 <-> u bit
// 6 + r, wave bit w <-> u bit 6 + RB + w.  Registers 2i and 2i + 1 then hold
// two consecutive amplitudes, so each load or store moves 16 bytes per lane.
__host__ __device__ constexpr std::uint32_t m0_u(std::uint32_t lane, std::uint32_t wave, std::uint32_t j, int rb)
{
    return (j & 1u) | (lane << 1) | ((j >> 1) << 7) | (wave << (6 + rb));
}

// FP64 histogram partial of a tile over its local measured sites, from the
// registers in the M0 mapping.  The local measured u bits split into lane
// (ml), register (mr) and wave (mw) bits; a bin is pext(u, loc_u), the wave
// bits highest.  Weights |psi|^2 are formed in FP64 and summed over the
// unmeasured register bits by in-register folds, over the unmeasured lane
// bits by butterfly exchanges (a + b in both lanes, exactly commutative) and
// over the waves in LDS in increasing wave order: a fixed order, so the
// partials are reproducible.  With cond, only amplitudes whose u bits under
// cdm equal cdv count.  red: LDS scratch of NT/64 * 2^(|ml| + |mr|) doubles.
// Every thread must call.
template <int NV, int NT>
__device__ __forceinline__ void tile_histogram(const float2 (&v)[NV], std::uint32_t loc_u, bool cond, std::uint32_t cdm,
                                               std::uint32_t cdv, int lane, int wave, double *red, double *out)
{
    constexpr int RB = NV == 32 ? 5 : 6;
    const std::uint32_t ml = (loc_u >> 1) & 63u;
    const std::uint32_t mr = (loc_u & 1u) | (((loc_u >> 7) & (NV / 2 - 1u)) << 1);
    const std::uint32_t mw = loc_u >> (6 + RB);
    const std::uint32_t nw_mask = (1u << (6 + RB)) - 1u;  // u bits below the wave bits
    double w[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j)
    {
        const double xr = v[j].x, xi = v[j].y;
        double q = fma(xr, xr, xi * xi);
        if (cond && ((m0_u(lane, wave, j, RB) & cdm) != cdv)) q = 0.0;
        w[j] = q;
    }
    // unmeasured register bits: fold j | 1 << r into j
#pragma unroll
    for (int r = 0; r < RB; ++r)
        if (!((mr >> r) & 1u))
        {
#pragma unroll
            for (int j = 0; j < NV; ++j)
                if (!(j & (1 << r))) w[j] += w[j | (1 << r)];
        }
    // unmeasured lane bits: butterfly over the surviving registers (one
    // uniform test per register, then its exchanges in increasing lane bit)
#pragma unroll
    for (int j = 0; j < NV; ++j)
        if ((static_cast<std::uint32_t>(j) & ~mr) == 0u)
        {
#pragma unroll
            for (int b = 0; b < 6; ++b)
                if (!((ml >> b) & 1u)) w[j] += __shfl_xor(w[j], 1 << b);
        }
    const int nlow = __popc(loc_u & nw_mask);
    const std::uint32_t stride = 1u << nlow;
    __syncthreads();  // the tile buffer is reused as scratch
    if ((static_cast<std::uint32_t>(lane) & ~ml) == 0u)
    {
#pragma unroll
        for (int j = 0; j < NV; ++j)
            if ((static_cast<std::uint32_t>(j) & ~mr) == 0u)
            {
                const std::uint32_t low = static_cast<std::uint32_t>(pext64(m0_u(lane, 0, j, RB), loc_u & nw_mask));
                red[static_cast<std::uint32_t>(wave) * stride + low] = w[j];
            }
    }
    __syncthreads();
    constexpr int NW = NT / 64;
    const std::uint32_t bins = stride << __popc(mw);
    for (std::uint32_t bidx = threadIdx.x; bidx < bins; bidx += NT)
    {
        const std::uint32_t low = bidx & (stride - 1u), wpart = bidx >> nlow;
        double acc = 0.0;
        for (int ww = 0; ww < NW; ++ww)
            if ((static_cast<std::uint32_t>(ww) & mw) == static_cast<std::uint32_t>(pdep64(wpart, mw)))
                acc += red[static_cast<std::uint32_t>(ww) * stride + low];
        out[bidx] = acc;
    }
}

// Finalize one chunk of a layer's measurements for one trajectory (one
// workgroup, after the histogram sweep): fixed-order reduction of the tile
// partials into the chunk's joint histogram h over its m sites (bin bit j <->
// j-th chunk site, ascending), a tree of marginals M_j[x] = sum of h over bins
// whose low j bits are x, then the sequential Born walk: site j's conditional
// P(1) = M_{j+1}[prefix | 1<<j] / M_j[prefix], outcome 1 iff u_j < P(1)
// (cpp/include/mipt/free_energy_measure.hpp:59), near-threshold flags, and the
// collapse record (mask, outcomes, 1/sqrt(raw joint weight)) for the next sweep.
struct Geom
{
    std::uint64_t rest_mask;
    int ntiles, k, c, s;
};

__device__ void finalize_chunk(const DevArena &a, int traj, int tl, int t, int chunk, const Geom g, double *lds_d)
{
    const TrajDesc d = a.desc[traj];
    const std::uint32_t fl = a.flags[d.layer_off + t];
    int sites[MAX_HIST_BITS];
    int m = 0, seen = 0;
    for (int q = 0; q < d.L; ++q)
    {
        if (!((fl >> q) & 1u)) continue;
        if (seen >= chunk * MAX_HIST_BITS && seen < (chunk + 1) * MAX_HIST_BITS) sites[m++] = q;
        ++seen;
    }
    if (m == 0)
    {
        if (threadIdx.x == 0 && chunk == 0)
        {
            a.coll[d.layer_off + t] = Collapse{0, 0, 1.0, 1.0};
            a.outcomes[d.layer_off + t] = 0;
            a.flagged[d.layer_off + t] = 0;
        }
        return;
    }
    // global-bin bits of local sites (ascending u bit = ascending site) and of tile sites
    std::uint32_t gloc = 0, gnon = 0;
    std::uint64_t non_phys = 0;
    for (int j = 0; j < m; ++j)
    {
        const int q = sites[j];
        const bool local = q < g.c || (q >= g.s && q < g.s + (g.k - g.c));
        if (local) gloc |= 1u << j;
        else
        {
            gnon |= 1u << j;
            non_phys |= std::uint64_t{1} << q;
        }
    }
    const std::uint32_t bins = 1u << m;
    double *M = lds_d;  // level j at M[2^j .. 2^(j+1)), level m = h
    double *h = M + bins;
    // tiles contributing to a global bin are those whose tile-index bits at the
    // chunk's tile sites equal the bin's tile part: enumerate them in
    // increasing tile order (deterministic), each partial read exactly once
    const std::uint32_t ntile_bits = static_cast<std::uint32_t>(__popcll(g.rest_mask));
    const std::uint32_t tmask_all = (ntile_bits >= 32) ? 0xffffffffu : ((1u << ntile_bits) - 1u);
    const std::uint32_t nmask = static_cast<std::uint32_t>(pext64(non_phys, g.rest_mask));  // tile-index bits
    const std::uint32_t fmask = tmask_all & ~nmask;
    const std::uint32_t per_bin = 1u << __popc(fmask);
    const double *part = a.partials + static_cast<std::size_t>(tl) * g.ntiles * HSTRIDE;
    for (std::uint32_t x = threadIdx.x; x < bins; x += blockDim.x)
    {
        const std::uint32_t bloc = static_cast<std::uint32_t>(pext64(x, gloc));
        const std::uint32_t tsel = static_cast<std::uint32_t>(pdep64(pext64(x, gnon), nmask));
        double acc = 0.0;
        std::uint32_t r = 0;
#pragma unroll 8
        for (std::uint32_t e = 0; e < per_bin; ++e)
        {
            acc += part[static_cast<std::size_t>(tsel | r) * HSTRIDE + bloc];
            r = next_submask(r, fmask);
        }
        h[x] = acc;
    }
    __syncthreads();
    for (int j = m - 1; j >= 0; --j)
    {
        const std::uint32_t n = 1u << j;
        for (std::uint32_t x = threadIdx.x; x < n; x += blockDim.x) M[n + x] = M[2 * n + x] + M[2 * n + x + n];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double *u = a.uniforms + d.unif_off + static_cast<std::size_t>(t) * d.L;
    std::uint32_t prefix = 0;
    int status = 0;
    std::uint32_t outc = (chunk == 0) ? 0u : a.outcomes[d.layer_off + t];
    std::uint32_t flg = (chunk == 0) ? 0u : a.flagged[d.layer_off + t];
    Collapse cc = (chunk == 0) ? Collapse{0, 0, 1.0, 1.0} : a.coll[d.layer_off + t];
    for (int j = 0; j < m; ++j)
    {
        const std::uint32_t n = 1u << j;
        const double w = M[n + prefix];           // marginal of the decided prefix
        const double w1 = M[2 * n + (prefix | n)];  // ... with site j = 1
        if (!(w > 0.0)) status = E_NORM;
        double p1 = (w > 0.0) ? w1 / w : 0.0;
        p1 = p1 < 0.0 ? 0.0 : (p1 > 1.0 ? 1.0 : p1);
        const int q = sites[j];
        const double uj = u[q];
        a.p1[d.unif_off + static_cast<std::size_t>(t) * d.L + q] = p1;
        cc.mask |= std::uint64_t{1} << q;
        if (uj < p1)
        {
            prefix |= n;
            cc.val |= std::uint64_t{1} << q;
            outc |= 1u << q;
        }
        if (fabs(uj - p1) < 1e-6) flg |= 1u << q;
    }
    const double total = M[1];
    const double w = h[prefix];
    if (status != 0 || !(w > 0.0) || !(total > 0.0))
    {
        a.status[traj] = E_NORM;
        cc.scale = 1.0;
    }
    else
    {
        cc.weight = w;
        cc.scale = 1.0 / sqrt(w);
        a.surprisal[traj] += -log(w / total);
    }
    a.coll[d.layer_off + t] = cc;
    a.outcomes[d.layer_off + t] = outc;
    a.flagged[d.layer_off + t] = flg;
}

// Wave-uniform context of one tile of a sweep: what its load, stages,
// histogram and store need (kept small: two are live at once).
struct TileCtx
{
    int id, tl, tile, traj, L, nq;
    std::uint64_t B;              // tile base (physical index)
    float2 *state;                // the trajectory's state
    std::uint64_t goff;           // the layer's first gate (complex64 element offset)
    std::uint32_t cml, cvl;       // previous collapse on local bits (physical)
    float scale;                  // its renormalisation
    bool zero, gen_init, tile_out;
    std::uint32_t flags;          // the layer's measured sites
    std::uint32_t cdm, cdv;       // decided outcomes on local bits (u bits; overflow-chunk histograms)
    std::uint64_t goff2;          // the next layer's first gate (fused launches)
};

// Read-only arrays of a sweep, from kernel arguments declared __restrict__ so
// that their wave-uniform reads become scalar loads: scalar loads complete on
// their own counter, so reading them never waits for a tile prefetch.
struct SweepIn
{
    const int *traj_list;
    const TrajDesc *desc;
    const std::uint32_t *flags;
    const Collapse *coll;
    const int *init;
    const float2 *gates;
};

// Context of tile `id`; false when the tile has nothing to do in this sweep
// (a collapsed-out tile the next sweep re-zeroes, or an overflow-chunk
// histogram of a trajectory without that chunk).
__device__ __forceinline__ bool tile_context(const PassArgs &args, const int *__restrict__ traj_list,
                                             const TrajDesc *__restrict__ desc, const std::uint32_t *__restrict__ flags,
                                             const Collapse *__restrict__ coll, const int *__restrict__ init,
                                             const float2 *__restrict__ gates, int id, bool aux, TileCtx &x)
{
    const PassPlan &pl = args.plan;
    const DevArena &a = args.a;
    const int t = args.layer;
    x.id = id;
    x.tl = id / pl.ntiles;
    x.tile = id - x.tl * pl.ntiles;
    // Every record of the tile is read up front, unconditionally, in three
    // dependent rounds of scalar loads (trajectory; its descriptor and init
    // flag; its layer's flags and collapse records): the tile's global loads
    // wait for no more round trips than these.
    const int entry = traj_list[x.tl];
    const bool pending = entry < 0;
    x.traj = pending ? ~entry : entry;
    const TrajDesc d = desc[x.traj];
    const int init_v = init[x.traj];
    const std::uint32_t lt = d.layer_off + static_cast<std::uint32_t>(t);
    const std::uint32_t fl = flags[lt];
    const Collapse cprev = coll[t > 0 ? lt - 1 : lt];
    const Collapse ccur = coll[lt];
    x.flags = fl;
    if (aux && pl.cond && __popc(x.flags) <= pl.chunk * MAX_HIST_BITS) return false;
    x.goff2 = (d.gate_off + gate_prefix(d, t + 1)) * 16;
    x.L = d.L;
    x.nq = d.nq;
    x.B = pdep64(static_cast<std::uint64_t>(x.tile), pl.rest_mask);
    x.state = a.states + d.state_off;
    x.goff = (d.gate_off + gate_prefix(d, t)) * 16;
    // the previous layer's collapse, applied at load (collapse), or, where
    // the layer's first sweep skipped tiles of its zero branch, that branch
    // on the pmask sites (untouched since, so still zero by definition);
    // selects rather than branches, so no load waits behind a branch
    const bool use_c = (pl.collapse || (pl.pend && pending)) && t > 0;
    const bool use_p = !use_c && pl.pmask != 0 && t > 0;
    const std::uint64_t pm = use_c ? ~std::uint64_t{0} : (use_p ? pl.pmask : 0);
    const std::uint64_t cm = cprev.mask & pm, cv = cprev.val & cm;
    x.scale = use_c ? static_cast<float>(cprev.scale) : 1.f;
    x.zero = (x.B & cm) != (cv & cm & pl.rest_mask);
    if (pl.zskip && x.zero) return false;
    x.cml = static_cast<std::uint32_t>(cm & ~pl.rest_mask);
    x.cvl = static_cast<std::uint32_t>(cv & ~pl.rest_mask);
    x.gen_init = pl.init && init_v;
    x.tile_out = false;
    x.cdm = x.cdv = 0;
    if (aux && pl.cond)
    {
        x.tile_out = (x.B & ccur.mask) != (ccur.val & ccur.mask & pl.rest_mask);
        const std::uint32_t lowmask = (1u << pl.c) - 1u, wmask = (1u << (pl.k - pl.c)) - 1u;
        x.cdm = static_cast<std::uint32_t>((ccur.mask & lowmask) | (((ccur.mask >> pl.s) & wmask) << pl.c));
        x.cdv = static_cast<std::uint32_t>((ccur.val & lowmask) | (((ccur.val >> pl.s) & wmask) << pl.c));
    }
    return true;
}

// Issues the loads of tile x in the M0 mapping (element (lane, wave, j) at
// in-tile offset toff + j << jshift); the collapsed-out branch of the
// previous layer's measurement reads as 0 (its scale is applied when the tile
// is processed).  A tile of an initial state is generated when processed.
template <int NV>
__device__ __forceinline__ void tile_load(const TileCtx &x, bool have, const float2 *any_state, std::uint32_t toff,
                                          std::uint32_t jshift, float2 (&v)[NV])
{
    // One branch-free sequence of 16-byte buffer loads for every case (a
    // merge of differently loaded registers would force a wait): a lane's
    // pair is fetched only when the tile exists, is neither all zero nor an
    // initial state, and the pair is not wholly on the previous collapse's
    // zero branch; every other lane takes an offset past the resource's
    // range, which the hardware answers with 0 and no fetch (states up to
    // 2^27 amplitudes, so in-range offsets stay below 2^31).  When qubit 0
    // itself was measured, one amplitude of each pair is zeroed when the tile
    // is processed.
    const bool fetch = have && !x.zero && !x.gen_init;
    const char *tp = fetch ? reinterpret_cast<const char *>(x.state + x.B) : reinterpret_cast<const char *>(any_state);
    const std::uint32_t range = fetch ? static_cast<std::uint32_t>(((std::uint64_t{1} << x.nq) - x.B) << 3) : 0u;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(tp), static_cast<short>(0),
                                                                        static_cast<int>(range), 0x00020000);
    const std::uint32_t m = fetch ? (x.cml & ~1u) : 0u, val = fetch ? (x.cvl & ~1u) : 0u;
#pragma unroll
    for (int i = 0; i < NV / 2; ++i)
    {
        const std::uint32_t o = toff + (static_cast<std::uint32_t>(i) << jshift);
        const std::uint32_t off = (fetch && (o & m) == val) ? (o << 3) : 0x80000000u;
        // nontemporal (cache policy nt): a sweep reads each tile once
        const float4 q = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 2));
        v[2 * i] = make_float2(q.x, q.y);
        v[2 * i + 1] = make_float2(q.z, q.w);
    }
}

// The thread's index in its workgroup, as a value the compiler cannot tie to
// an earlier computation: per-lane terms derived from it are recomputed where
// they are needed instead of held in registers across the LDS stages.
__device__ __forceinline__ std::uint32_t fresh_tid()
{
    std::uint32_t t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// Stage structure of a sweep plan as a compile-time code: bits 0-1 the
// stage count (1..3); per stage si three bits from bit 2 + 3 si: an exchange
// starts it (only stage 0 may work in the load mapping), bond[0] present,
// bond[1] present.  The bond of qubit 0, when the sweep has one, is the last
// stage's bond[0] (its form is the kernel's ZP).  Sweep kernels with a code
// run exactly that structure as straight-line code: no run-time test of a
// stage's exchange or bonds, whose two outcomes would leave the tile's
// registers paired differently (the bond on register bits (3,4) pairs
// registers j, j|8 in its outputs, the one on (1,2) j, j|2) and force copies
// back to one layout where they meet.
__host__ __device__ constexpr int pat_nst(int c) { return c & 3; }
__host__ __device__ constexpr bool pat_lds(int c, int si) { return (c >> (2 + 3 * si)) & 1; }
__host__ __device__ constexpr bool pat_b0(int c, int si) { return (c >> (3 + 3 * si)) & 1; }
__host__ __device__ constexpr bool pat_b1(int c, int si) { return (c >> (4 + 3 * si)) & 1; }

// The compiled-structure sweep kernels (matrix-core bonds, 2^13 tiles of 256
// threads): zp 0..2 a sweep of that qubit-0 form with structure pat; zp 4..6
// a fused launch, base structure pat (form 0), continuation pat2 of form
// zp - 4.  They cover every sweep and fused pair of the L=20 ring and chain
// plans, of the L=24 and L=26 rings and of the 4 x 5 grid (every plan
// variant); other sweeps run the stage loop that reads the structure at run
// time.
struct CompiledSweep
{
    int zp, pat, pat2;
};
constexpr CompiledSweep kCompiledSweeps[] = {
    {0, 1019, 0}, {1, 1023, 0}, {2, 1919, 0}, {5, 1019, 1023}, {6, 1019, 1919},                  // L=20
    {0, 250, 0},  {0, 254, 0},  {1, 254, 0},  {1, 891, 0},     {1, 1915, 0},   {2, 254, 0},  {2, 895, 0},  // L=24, 26
    {4, 250, 254}, {6, 250, 254}, {6, 250, 895}, {4, 1019, 254},
    {0, 29, 0},   {0, 242, 0},  {1, 110, 0},  {4, 242, 254},   {4, 1019, 29},  {5, 254, 1023},  // 4 x 5 grid
};
constexpr int kNumCompiledSweeps = static_cast<int>(sizeof(kCompiledSweeps) / sizeof(kCompiledSweeps[0]));

// Processes one tile whose amplitudes are in v (M0 mapping): the initial
// state or the previous layer's collapse scale, the stages of the sweep plan,
// back to M0, the store, and the FP64 histogram partial.  ZP 1 / 2: the last
// stage holds the sweep's bond of qubit 0 in that form and the others none;
// ZP 0: no stage has one; ZP 3: any stage may, in either form.
template <int RB, int NT, int AUX, int MX, int ZP, int PAT, int PAT2>
__device__ __forceinline__ void process_tile(const PassArgs &args, const float2 *__restrict__ gates, const TileCtx &cur,
                                             float2 (&v)[1 << RB], std::uint32_t jshift,
                                             const std::uint32_t (&Q3m0)[6])
{
    const std::uint32_t lowmask = (1u << args.plan.c) - 1u;
    // per-lane terms of the M0 mapping, from a fresh thread index at each use
    auto m0_terms = [&](std::uint32_t &toff, std::uint32_t &T3m0, int &lane, int &wave) {
        const std::uint32_t t = fresh_tid();
        lane = static_cast<int>(t & 63u);
        wave = static_cast<int>(t >> 6);
        const std::uint32_t u = m0_u(t & 63u, t >> 6, 0, RB);
        toff = (u & lowmask) | ((u >> args.plan.c) << args.plan.s);
        T3m0 = swz(u >> 1) << 3;
    };
    constexpr int NV = 1 << RB;
    constexpr int WB = (NT == 256) ? 2 : (NT == 128 ? 1 : 0);
    extern __shared__ float2 lds[];
    const PassPlan &pl = args.plan;
    const DevArena &a = args.a;
    const int k = pl.k;
    const int c = pl.c, s = pl.s;
    char *ldsb = reinterpret_cast<char *>(lds);
    // ---- initial state, or the collapse scale of the current tile
    std::uint32_t toff, T3m0;
    int lane, wave;
    m0_terms(toff, T3m0, lane, wave);
    if (cur.gen_init)
    {
        const TrajDesc d = a.desc[cur.traj];
#pragma unroll
        for (int j = 0; j < NV; ++j)
            v[j] = initial_amp(d, cur.B | (toff + ((static_cast<std::uint32_t>(j) >> 1) << jshift) + (j & 1)));
    }
    else if (cur.scale != 1.f || (cur.cml & 1u))
    {
        // qubit 0 measured: one amplitude of each pair is on the zero branch
#pragma unroll
        for (int j = 0; j < NV; ++j)
        {
            const bool off_branch = (cur.cml & 1u) && ((static_cast<std::uint32_t>(j) & 1u) != (cur.cvl & 1u));
            const float f = off_branch ? 0.f : cur.scale;
            v[j].x *= f;
            v[j].y *= f;
        }
    }

    // ---- stages
    const float2 *__restrict__ gl = gates + cur.goff;
    std::uint32_t T3 = T3m0;
    std::uint32_t Q3[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) Q3[r] = Q3m0[r];
    bool in_m0 = true;
    bool lds_dirty = false;  // a previous LDS read must finish before the next write
    // matrix-core coefficient table, after the half-tile buffer; its writes
    // are published by the first exchange's barrier (or one of its own)
    const float *mxtab = reinterpret_cast<const float *>(ldsb + (std::size_t{8} << (k - 1)));
    bool tab_ready = MX == 0;
    constexpr bool FUSED = ZP >= 4;  // ZP 4..6: the next layer's first sweep follows, its qubit-0 form ZP - 4
    if constexpr (MX > 0 && AUX == 0)
    {
        fill_mx_table<NT>(pl, gates + cur.goff, const_cast<float *>(mxtab));
        if constexpr (FUSED)
            fill_mx_table<NT>(args.next, gates + cur.goff2,
                              const_cast<float *>(mxtab) + pl.nstages * (2 * MX_SLOT_BYTES / 4));
    }
    // moves the tile into stage sp's mapping (no move for an M0 stage: the
    // first of a plan, and a fused continuation starts with an exchange)
    auto exchange_to = [&](const StagePlan &sp) {
        const std::uint32_t tt = fresh_tid();
        const std::uint32_t T3n =
            thread_term<WB>(sp.lane_u, sp.wave_u, static_cast<int>(tt & 63u), static_cast<int>(tt >> 6)) << 3;
        std::uint32_t Q3n[6];
        Q3n[0] = 0;
#pragma unroll
        for (int r = 1; r < 6; ++r) Q3n[r] = swz((1u << sp.reg_u[r]) >> 1) << 3;
        lds_exchange<NV>(ldsb, v, T3, Q3, T3n, Q3n, lds_dirty);
        T3 = T3n;
#pragma unroll
        for (int r = 0; r < 6; ++r) Q3[r] = Q3n[r];
        in_m0 = false;
    };
    auto enter_stage = [&](const StagePlan &sp) {
        if (!tab_ready && !sp.use_lds) __syncthreads();
        tab_ready = true;
        if (sp.use_lds) exchange_to(sp);
    };
    // the stages of plan P, its qubit-0 form Z (the stage type of ztag) fixed
    // at compile time, gates glx, coefficient-table slots from stage slot0
    auto run_stages = [&](const PassPlan &P, auto ztag, auto ptag, const float2 *glx, int slot0) {
        constexpr int Z = decltype(ztag)::value;
        constexpr int C = decltype(ptag)::value;
        if constexpr (C != 0 && AUX == 0)
        {
            static_assert(Z <= 2 && MX != 0, "compiled stage structure: forms 0..2 on the matrix cores");
            static_assert(pat_nst(C) >= 1 && (Z == 0 || pat_b0(C, pat_nst(C) - 1)), "qubit-0 bond is the last bond[0]");
            auto stage = [&](auto sidx) {
                constexpr int si = decltype(sidx)::value;
                if constexpr (si < pat_nst(C))
                {
                    constexpr int zf = si == pat_nst(C) - 1 ? Z : 0;
                    const StagePlan &sp = P.st[si];
                    if constexpr (pat_lds(C, si)) exchange_to(sp);
                    else if (!tab_ready) __syncthreads();
                    tab_ready = true;
                    const float *tab = mxtab + (slot0 + si) * (2 * MX_SLOT_BYTES / 4);
                    if constexpr (pat_b0(C, si))
                    {
                        const float2 *g0 = glx + 16 * sp.bond[0];
                        if constexpr (zf == 1) bond_on_regs<NV, 0, 1, MX>(v, g0, tab, lane);
                        else if constexpr (zf == 2) bond_on_regs<NV, 1, 0, MX>(v, g0, tab, lane);
                        else bond_on_regs<NV, 1, 2, MX>(v, g0, tab, lane);
                    }
                    if constexpr (pat_b1(C, si))
                        bond_on_regs<NV, 3, 4, MX>(v, glx + 16 * sp.bond[1], tab + MX_SLOT_BYTES / 4, lane);
                }
            };
            stage(std::integral_constant<int, 0>{});
            stage(std::integral_constant<int, 1>{});
            stage(std::integral_constant<int, 2>{});
            return;
        }
        constexpr bool zlast = Z == 1 || Z == 2;
        const int nloop = AUX ? 0 : P.nstages - (zlast ? 1 : 0);
        for (int si = 0; si < nloop; ++si)
        {
            const StagePlan &sp = P.st[si];
            enter_stage(sp);
            apply_stage_gates<NV, MX, Z == 3 ? 3 : 0>(v, sp, glx, mxtab + (slot0 + si) * (2 * MX_SLOT_BYTES / 4), lane);
        }
        if constexpr (zlast && !AUX)
        {
            const int si = P.nstages - 1;
            const StagePlan &sp = P.st[si];
            enter_stage(sp);
            apply_stage_gates<NV, MX, Z>(v, sp, glx, mxtab + (slot0 + si) * (2 * MX_SLOT_BYTES / 4), lane);
        }
    };
    run_stages(pl, std::integral_constant<int, FUSED ? 0 : ZP>{}, std::integral_constant<int, PAT>{}, gl, 0);

    // back to the load mapping, the store, the histogram partial
    auto finish = [&]() {
        // ---- the chunk's measured sites that are local to this tile (u bits)
        std::uint32_t loc_u = 0;
        // a layer without measured sites needs no partials (its finalize
        // records the identity collapse without reading them)
        const bool do_hist = pl.hist && !(pl.chunk == 0 && cur.flags == 0u);
        if (do_hist)
        {
            int seen = 0;
            for (int q = 0; q < cur.L; ++q)
            {
                if (!((cur.flags >> q) & 1u)) continue;
                if (seen >= pl.chunk * MAX_HIST_BITS && seen < (pl.chunk + 1) * MAX_HIST_BITS)
                {
                    if (q < c) loc_u |= 1u << q;
                    else if (q >= s && q < s + (k - c)) loc_u |= 1u << (q - s + c);
                }
                ++seen;
            }
        }

        // ---- back to M0 through LDS
        m0_terms(toff, T3m0, lane, wave);
        if (!in_m0) lds_exchange<NV>(ldsb, v, T3, Q3, T3m0, Q3m0, lds_dirty);

        // ---- store (M0); before the histogram, so the tile's registers are
        // free again as its FP64 weights are formed.  Nontemporal, as the
        // loads: the tile is not read again before the next sweep, after
        // every other tile of the launch (with both, the sweeps measured 2%
        // faster)
        if (pl.store)
        {
            char *tp = reinterpret_cast<char *>(cur.state + cur.B);
#pragma unroll
            for (int i = 0; i < NV / 2; ++i)
                __builtin_nontemporal_store(
                    floatx4_t{v[2 * i].x, v[2 * i].y, v[2 * i + 1].x, v[2 * i + 1].y},
                    reinterpret_cast<floatx4_t *>(tp + ((toff + (static_cast<std::uint32_t>(i) << jshift)) << 3)));
        }
        // ---- FP64 histogram partial of this tile over the chunk's local sites
        if (do_hist)
        {
            double *out = a.partials + (static_cast<std::size_t>(pl.pbase + cur.tl) * pl.ntiles + cur.tile) * HSTRIDE;
            if (cur.tile_out)
            {
                // the tile lies outside the decided outcomes: every bin is 0
                const std::uint32_t bins = 1u << __popc(loc_u);
                for (std::uint32_t bb = threadIdx.x; bb < bins; bb += NT) out[bb] = 0.0;
            }
            else
                tile_histogram<NV, NT>(v, loc_u, AUX && pl.cond, cur.cdm, cur.cdv, lane, wave,
                                       reinterpret_cast<double *>(lds), out);
        }
    };
    // a fused launch holds the trajectories that fuse this layer into the
    // next: the next layer's first sweep (same window; its bonds touch none
    // of this layer's measured sites, whose collapse waits) follows on the
    // tile in registers and LDS, and the histogram is taken after it
    if constexpr (FUSED)
        run_stages(args.next, std::integral_constant<int, ZP - 4>{}, std::integral_constant<int, PAT2>{}, gates + cur.goff2,
                   pl.nstages);
    finish();
}

// One sweep over the tiles of a group of trajectories.  AUX = 0: a gate sweep
// (the hot kernel); AUX = 1: a sweep without gates (the closing collapse, or
// the histogram of an extra chunk of a layer with more than MAX_HIST_BITS
// measured sites), kept as its own kernel so profiles separate the two.
// One workgroup per tile; with the half-tile LDS buffer four 2^13-amplitude
// workgroups share a CU (64-register tiles: one), whose loads overlap one
// another's LDS stages and bond arithmetic.
template <int RB, int NT, int AUX, int MX, int ZP, int PAT = 0, int PAT2 = 0>
__global__ void __launch_bounds__(NT, (RB == 6 ? 1 : 4)) pass_kernel(PassArgs args, const int *__restrict__ traj_list,
                                                                  const TrajDesc *__restrict__ desc,
                                                                  const std::uint32_t *__restrict__ flags,
                                                                  const Collapse *__restrict__ coll,
                                                                  const int *__restrict__ init,
                                                                  const float2 *__restrict__ gates)
{
    constexpr int NV = 1 << RB;
    static_assert(RB == 5 || RB == 6, "register tiles hold 32 or 64 amplitudes");
    const PassPlan &pl = args.plan;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int c = pl.c, s = pl.s;
    const std::uint32_t lowmask = (1u << c) - 1u;
    // M0 element offset of (lane, wave, register j): the thread part toff
    // (lane bits on u bits 1..6, wave bits above the registers) plus (j & 1)
    // + ((j >> 1) << jshift) (register bits 1.. on u bits 7..)
    auto phys = [&](std::uint32_t u) { return (u & lowmask) | ((u >> c) << s); };
    const std::uint32_t jshift = static_cast<std::uint32_t>(s + 7 - c);
    const std::uint32_t toff = phys(m0_u(static_cast<std::uint32_t>(lane), static_cast<std::uint32_t>(wave), 0, RB));
    // the M0 mapping's uniform LDS terms (its per-lane term is formed where used)
    std::uint32_t Q3m0[6];
    Q3m0[0] = 0;
#pragma unroll
    for (int r = 1; r < 6; ++r) Q3m0[r] = swz(m0_u(0, 0, 1u << r, RB) >> 1) << 3;

    const int total = args.ntraj * pl.ntiles;
    const SweepIn in{traj_list, desc, flags, coll, init, gates};
    TileCtx x;
    float2 v[NV];
    const int id = static_cast<int>(blockIdx.x);
    if (id < total && tile_context(args, in.traj_list, in.desc, in.flags, in.coll, in.init, in.gates, id, AUX != 0, x))
    {
        tile_load<NV>(x, true, args.a.states, toff, jshift, v);
        process_tile<RB, NT, AUX, MX, ZP, PAT, PAT2>(args, in.gates, x, v, jshift, Q3m0);
    }
}

// Which trajectories fuse layer t into layer t+1 (their last sweep of layer
// t and first of layer t+1 in one launch): those that measure at most
// FUSE_MAX_MEASURED sites in layer t, none of them bonded in layer t+1's
// first sweep.  A trajectory that measures more keeps the two sweeps: the
// first of them then skips the tiles of the collapse's zero branch and the
// second masks the rest of it, which moves fewer bytes than the fused pass
// (measured at L=20 and 26: fusing up to 0, 1, 2 or 10 measured sites gave
// the same rate within 0.5% at L=20, and 1.7% / 1% / 0.5% more than 10 at
// L=26).  Those bonds commute with layer t's projectors and
// leave the joint distribution of its measured sites unchanged, so the
// histogram is taken after them and the collapse waits for layer t+1's
// second sweep.  Per layer parity j = t mod period: ok[j] (the two sweeps
// share a window the fused kernel takes) and block[j] (layer t+1's first
// sweep's bonded sites).  The host applies the same rule to its copy of the
// flags to size the launches.
constexpr int FUSE_MAX_MEASURED = 1;
static_assert(FUSE_MAX_MEASURED <= MAX_HIST_BITS, "a fused launch takes one histogram chunk");

struct FuseRule
{
    std::uint64_t block[4];
    int ok[4];
    int period;
    int t1;  // end of the evolution range: its last layer fuses into nothing
};

__host__ __device__ inline bool fuses(const FuseRule &r, std::uint32_t flags, int t)
{
    const int j = t % r.period;
    return t + 1 < r.t1 && r.ok[j] && __builtin_popcount(flags) <= FUSE_MAX_MEASURED && (flags & r.block[j]) == 0;
}

// Layer fusion's trajectory order: for layer t = t0 + blockIdx.x, the
// group's trajectories (entries of list) that fuse layer t first, then the
// others, each in list order, into out[blockIdx.x * n + ...]; a trajectory
// that fused layer t-1 (t > t0) is entered as ~index (its collapse of layer
// t-1 is still to apply).  q: n bytes of LDS.
__global__ void __launch_bounds__(256) fuse_order_kernel(const int *__restrict__ list, int n,
                                                         const TrajDesc *__restrict__ desc,
                                                         const std::uint32_t *__restrict__ flags, int t0, FuseRule r,
                                                         int *__restrict__ out)
{
    extern __shared__ unsigned char q[];
    const int t = t0 + static_cast<int>(blockIdx.x);
    for (int j = threadIdx.x; j < n; j += 256)
    {
        const std::uint32_t off = desc[list[j]].layer_off;
        const bool now = fuses(r, flags[off + t], t);
        const bool before = t > t0 && fuses(r, flags[off + t - 1], t - 1);
        q[j] = static_cast<unsigned char>((now ? 1 : 0) | (before ? 2 : 0));
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int *o = out + static_cast<std::size_t>(blockIdx.x) * n;
        int w = 0;
        for (int j = 0; j < n; ++j)
            if (q[j] & 1) o[w++] = (q[j] & 2) ? ~list[j] : list[j];
        for (int j = 0; j < n; ++j)
            if (!(q[j] & 1)) o[w++] = (q[j] & 2) ? ~list[j] : list[j];
    }
}

// One workgroup per trajectory of the group: finalize_chunk after a
// histogram sweep (the kernel boundary publishes every tile's partials).
struct FinalizeArgs
{
    DevArena a;
    const int *traj_list;  // entries ~i (layer fusion's pending mark) name trajectory i
    int layer, chunk;
    Geom g;
    int pbase;  // partial record of the list's first trajectory
};

__global__ void __launch_bounds__(256) finalize_kernel(FinalizeArgs f)
{
    __shared__ double M[2 << MAX_HIST_BITS];
    const int entry = f.traj_list[blockIdx.x];
    finalize_chunk(f.a, entry < 0 ? ~entry : entry, f.pbase + static_cast<int>(blockIdx.x), f.layer, f.chunk, f.g, M);
}

// Marks the layer's collapse as applied (identity), after the closing sweep.
__global__ void reset_collapse_kernel(DevArena a, const int *__restrict__ traj_list, int n, int layer)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const TrajDesc d = a.desc[traj_list[i]];
    if (layer >= 0 && layer < d.layers) a.coll[d.layer_off + layer] = Collapse{0, 0, 1.0, 1.0};
}

__global__ void fill_int_kernel(int *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ============================================================== host side
namespace
{
struct Config
{
    int rb, nt, c;
    int mx = 0;  // bond arithmetic: 0 VALU, 1 VALU + matrix cores, 2 matrix cores
    // layer plan variant (qubit-block assignment): bit 0 puts a bond local to
    // several windows where it lands on a free register-slot pair (no LDS
    // exchange) before balancing counts; bit 1 spreads the overlaps of three
    // or more chain windows evenly instead of packing the slack into the last
    int plan = 0;
    int k() const { return 6 + rb + (nt == 256 ? 2 : nt == 128 ? 1 : 0); }
};

// Layer fusion: mipt_hip_set_layer_fusion, else MIPT_HIP_FUSE (1 on, the
// default; 0 off; any other value is refused).
std::atomic<int> g_fuse{-1};  // -1: not set by call (or set back)

int fuse_layers(bool &on)
{
    int v = g_fuse.load();
    if (v < 0)
    {
        static const int env = [] {
            const char *e = std::getenv("MIPT_HIP_FUSE");
            if (!e || !*e || std::strcmp(e, "1") == 0) return 1;
            if (std::strcmp(e, "0") == 0) return 0;
            return -1;
        }();
        v = env;
    }
    if (v < 0)
    {
        set_error("MIPT_HIP_FUSE must be 0 or 1");
        return E_ARG;
    }
    on = v == 1;
    return OK;
}

// Build the sweeps of one layer parity for a chain of L sites in nq qubits.
struct LayerPlan
{
    std::vector<PassPlan> passes;
};

// The M0 mapping as a stage plan (m0_u): register bit 0 <-> u bit 0,
// register bit r >= 1 <-> u bit 6 + r, lane bit b <-> u bit b + 1, waves above.
void set_m0(StagePlan &sp, int rb, int wb)
{
    for (int r = 0; r < 6; ++r) sp.reg_u[r] = static_cast<int>(r == 0 ? 0 : (r < rb ? 6 + r : 0));
    for (int b = 0; b < 6; ++b) sp.lane_u[b] = static_cast<int>(b + 1);
    for (int b = 0; b < 4; ++b) sp.wave_u[b] = static_cast<int>(b < wb ? 6 + rb + b : 0);
}

// Register bit of u bit x under M0 (-1: a lane or wave bit).
int m0_reg(int x, int rb)
{
    if (x == 0) return 0;
    if (x >= 7 && x < 6 + rb) return x - 6;
    return -1;
}

// Choose `count` lane bits (lane_u[0..count)) among `free` so that their
// low-nibble contributions under swz() cover distinct banks where possible
// (one bit per residue class mod 4 first); the chosen bits leave `free`.
void pick_lanes(std::vector<int> &freeb, int lane_u[6], int count)
{
    std::vector<int> chosen;
    for (int res = 0; res < 4 && static_cast<int>(chosen.size()) < count; ++res)
        for (std::size_t i = 0; i < freeb.size(); ++i)
            if ((freeb[i] & 3) == res && std::find(chosen.begin(), chosen.end(), freeb[i]) == chosen.end())
            {
                chosen.push_back(freeb[i]);
                break;
            }
    for (std::size_t i = 0; i < freeb.size() && static_cast<int>(chosen.size()) < count; ++i)
        if (std::find(chosen.begin(), chosen.end(), freeb[i]) == chosen.end()) chosen.push_back(freeb[i]);
    for (int b = 0; b < count; ++b) lane_u[b] = chosen[b];
    std::vector<int> rest;
    for (int x : freeb)
        if (std::find(chosen.begin(), chosen.end(), x) == chosen.end()) rest.push_back(x);
    freeb.swap(rest);
}

int build_layer_plan(int L, int nq, int closed, int gx, int layer, const Config &cf, LayerPlan &out)
{
    out.passes.clear();
    const int k = cf.k();
    const int c = cf.c;
    const int wbits = k - c;
    const int wb = cf.nt == 256 ? 2 : (cf.nt == 128 ? 1 : 0);
    if (nq <= k) return E_UNSUPPORTED;
    const int nb = geo_bond_count(L, gx, layer, closed);
    std::vector<int> b0(nb), b1(nb);
    for (int b = 0; b < nb; ++b) geo_bond(L, gx, layer, b, closed, b0[b], b1[b]);
    auto in_window = [&](int s, int q) { return q < c || (q >= s && q < s + wbits); };
    // windows over the physical bits c..nq-1: for a chain, consecutive
    // windows overlapping by one bit, so that every bond (i, i+1) lies in
    // some window; when a layout leaves a bond outside all of them (grid
    // columns), a greedy cover: the window start holding most bonds not yet
    // covered, until every bond is covered
    std::vector<int> starts;
    int st = c;
    while (true)
    {
        if (st + wbits >= nq)
        {
            starts.push_back(std::max(c, nq - wbits));
            break;
        }
        starts.push_back(st);
        st = st + wbits - 1;
    }
    auto covered_by_any = [&](int b) {
        for (int s0 : starts)
            if (in_window(s0, b0[b]) && in_window(s0, b1[b])) return true;
        return false;
    };
    if ((cf.plan & 2) && starts.size() >= 3)
    {
        const std::vector<int> packed = starts;
        const int W = static_cast<int>(starts.size());
        for (int i = 0; i < W; ++i)
            starts[i] = c + static_cast<int>(std::lround(static_cast<double>(i) * (nq - wbits - c) / (W - 1)));
        bool cov = true;
        for (int b = 0; b < nb && cov; ++b) cov = covered_by_any(b);
        if (!cov) starts = packed;
    }
    bool all = true;
    for (int b = 0; b < nb && all; ++b) all = covered_by_any(b);
    if (!all)
    {
        starts.clear();
        std::vector<char> done(nb, 0);
        int left = nb;
        while (left > 0)
        {
            int best_s = -1, best_n = 0;
            for (int s0 = c; s0 + wbits <= nq; ++s0)
            {
                int n_here = 0;
                for (int b = 0; b < nb; ++b) n_here += !done[b] && in_window(s0, b0[b]) && in_window(s0, b1[b]);
                if (n_here > best_n)
                {
                    best_n = n_here;
                    best_s = s0;
                }
            }
            if (best_s < 0)
            {
                set_error("tile plan: a bond of layer %d is local to no window (L=%d k=%d)", layer, L, k);
                return E_UNSUPPORTED;
            }
            starts.push_back(best_s);
            for (int b = 0; b < nb; ++b)
                if (!done[b] && in_window(best_s, b0[b]) && in_window(best_s, b1[b]))
                {
                    done[b] = 1;
                    --left;
                }
        }
    }
    auto local = [&](int p, int q) { return in_window(starts[p], q); };
    std::vector<std::vector<int>> assigned(starts.size());
    // register-slot pair a bond takes in window p without an LDS stage (0: bits
    // (1,2), 1: bits (3,4)), or -1
    std::vector<std::array<int, 2>> slot_used(starts.size(), std::array<int, 2>{0, 0});
    auto m0_slot_of = [&](int p, int b) {
        const int s0 = starts[p];
        auto uq = [&](int q) { return q < c ? q : q - s0 + c; };
        const int ra = m0_reg(uq(b0[b]), cf.rb), rc = m0_reg(uq(b1[b]), cf.rb);
        return (ra == 1 && rc == 2) ? 0 : ((ra == 3 && rc == 4) ? 1 : -1);
    };
    auto take = [&](int p, int b) {
        assigned[p].push_back(b);
        const int sl = m0_slot_of(p, b);
        if (sl >= 0) slot_used[p][sl] = 1;
    };
    // bonds local to exactly one window first, then balance the rest
    std::vector<int> flexible;
    for (int b = 0; b < nb; ++b)
    {
        std::vector<int> cand;
        for (std::size_t p = 0; p < starts.size(); ++p)
            if (local(static_cast<int>(p), b0[b]) && local(static_cast<int>(p), b1[b])) cand.push_back(static_cast<int>(p));
        if (cand.empty())
        {
            set_error("tile plan: bond (%d,%d) is local to no window (L=%d k=%d)", b0[b], b1[b], L, k);
            return E_UNSUPPORTED;
        }
        if (cand.size() == 1) take(cand[0], b);
        else flexible.push_back(b);
    }
    for (int b : flexible)
    {
        int best = -1;
        if (cf.plan & 1)
            for (std::size_t p = 0; p < starts.size() && best < 0; ++p)
            {
                const int pi = static_cast<int>(p);
                const int sl = (local(pi, b0[b]) && local(pi, b1[b])) ? m0_slot_of(pi, b) : -1;
                if (sl >= 0 && !slot_used[p][sl]) best = pi;
            }
        if (best < 0)
            for (std::size_t p = 0; p < starts.size(); ++p)
                if (local(static_cast<int>(p), b0[b]) && local(static_cast<int>(p), b1[b]) &&
                    (best < 0 || assigned[p].size() < assigned[best].size()))
                    best = static_cast<int>(p);
        take(best, b);
    }
    for (std::size_t p = 0; p < starts.size(); ++p)
    {
        PassPlan pp{};
        const int s = starts[p];
        pp.k = k;
        pp.c = c;
        pp.s = s;
        std::uint64_t local_mask = (std::uint64_t{1} << c) - 1;
        local_mask |= ((std::uint64_t{1} << wbits) - 1) << s;
        pp.rest_mask = ((std::uint64_t{1} << nq) - 1) & ~local_mask;
        pp.ntiles = 1 << (nq - k);
        pp.bonded = 0;
        for (int b : assigned[p]) pp.bonded |= (std::uint64_t{1} << b0[b]) | (std::uint64_t{1} << b1[b]);
        auto u_of = [&](int q) { return q < c ? q : q - s + c; };
        // M0 stage: bonds already on M0 register slots (1,2) / (3,4) (no LDS
        // exchange); u bit 0 is register bit 0 in every mapping
        std::vector<int> m0_bonds, other;
        int m0_slot[2] = {-1, -1};
        for (int b : assigned[p])
        {
            const int ra = m0_reg(u_of(b0[b]), cf.rb), rc = m0_reg(u_of(b1[b]), cf.rb);
            if (ra == 1 && rc == 2 && m0_slot[0] < 0) m0_slot[0] = b;
            else if (ra == 3 && rc == 4 && m0_slot[1] < 0) m0_slot[1] = b;
            else other.push_back(b);
        }
        std::vector<StagePlan> stages;
        if (m0_slot[0] >= 0 || m0_slot[1] >= 0)
        {
            StagePlan sp{};
            set_m0(sp, cf.rb, wb);
            sp.use_lds = 0;
            sp.zpair = 0;
            sp.bond[0] = m0_slot[0];
            sp.bond[1] = m0_slot[1];
            sp.nb = (m0_slot[0] >= 0) + (m0_slot[1] >= 0);
            stages.push_back(sp);
        }
        // LDS stages: two bonds each -- a bond of qubit 0 on bits (0,1) or
        // (1,0), any other on (1,2), and the second on (3,4) -- with u bit 0
        // kept on register bit 0
        std::size_t i = 0;
        while (i < other.size())
        {
            StagePlan sp{};
            sp.use_lds = 1;
            sp.bond[0] = sp.bond[1] = -1;
            sp.zpair = 0;
            for (int r = 0; r < 6; ++r) sp.reg_u[r] = -1;
            sp.reg_u[0] = 0;
            std::vector<int> used{0};
            int nbs = 0;
            // the first bond of the stage takes slot 0 (with qubit 0 if it has it)
            for (int slot = 0; slot < 2 && i < other.size(); ++slot)
            {
                const int b = other[i];
                const int ua = u_of(b0[b]), ub = u_of(b1[b]);
                const bool has0 = ua == 0 || ub == 0;
                if (slot == 1 && has0) break;  // a bond of qubit 0 needs slot 0
                ++i;
                ++nbs;
                if (slot == 0)
                {
                    sp.bond[0] = b;
                    if (has0)
                    {
                        sp.zpair = ua == 0 ? 1 : 2;  // q0 on bit 0 : q1 on bit 0
                        sp.reg_u[1] = ua == 0 ? ub : ua;
                        used.push_back(sp.reg_u[1]);
                    }
                    else
                    {
                        sp.reg_u[1] = ua;
                        sp.reg_u[2] = ub;
                        used.push_back(ua);
                        used.push_back(ub);
                    }
                }
                else
                {
                    sp.bond[1] = b;
                    sp.reg_u[3] = ua;
                    sp.reg_u[4] = ub;
                    used.push_back(ua);
                    used.push_back(ub);
                }
            }
            sp.nb = nbs;
            std::vector<int> freeb;
            for (int u = 0; u < k; ++u)
                if (std::find(used.begin(), used.end(), u) == used.end()) freeb.push_back(u);
            // fill the remaining register bits first from high u bits (keeps
            // lanes on bank-friendly bits), then lanes, then waves
            for (int r = 1; r < cf.rb; ++r)
                if (sp.reg_u[r] < 0)
                {
                    sp.reg_u[r] = freeb.back();
                    freeb.pop_back();
                }
            for (int rr = cf.rb; rr < 6; ++rr) sp.reg_u[rr] = 0;
            pick_lanes(freeb, sp.lane_u, 6);
            for (int b = 0; b < 4; ++b) sp.wave_u[b] = static_cast<int>(b < wb ? freeb[b] : 0);
            stages.push_back(sp);
        }
        if (stages.size() > 6)
        {
            set_error("tile plan: too many stages");
            return E_UNSUPPORTED;
        }
        // the stage with the bond of qubit 0 (at most one: a layer's bonds are
        // disjoint) goes last, where the sweep kernel applies it in a form
        // fixed at compile time; the bonds of one sweep commute
        std::stable_partition(stages.begin(), stages.end(), [](const StagePlan &sp) { return sp.zpair == 0; });
        pp.nstages = static_cast<int>(stages.size());
        for (std::size_t i = 0; i < stages.size(); ++i) pp.st[i] = stages[i];
        pp.store = 1;
        out.passes.push_back(pp);
    }
    return OK;
}
// References of a spec: count 0..4; a list needs probe_site = -1 and
// distinct sites in [0, L); at most 27 qubits in all.
bool probe_spec_ok(const mipt_hip_traj_spec &s)
{
    if (s.probe_count < 0 || s.probe_count > 4) return false;
    if (s.probe_count == 0) return s.L + (s.probe_site >= 0 ? 1 : 0) <= 27;
    if (s.probe_site >= 0 || s.L + s.probe_count > 27 || s.attach_stagger < 0) return false;
    const int first = s.attach_layer >= 0 ? s.attach_layer : s.p_zero_layers;
    if (s.attach_layer < -1) return false;
    if (s.probe_count > 1 && s.attach_stagger > 0 && first + (s.probe_count - 1) * s.attach_stagger >= s.layers)
        return false;
    if (first > 0 && first >= s.layers) return false;
    for (int j = 0; j < s.probe_count; ++j)
    {
        if (s.probe_sites[j] < 0 || s.probe_sites[j] >= s.L) return false;
        // one site can serve several references only at different attach layers
        for (int k = 0; k < j; ++k)
            if (s.probe_sites[k] == s.probe_sites[j] && s.attach_stagger == 0) return false;
    }
    return true;
}

} // namespace

struct Batch
{
    hipStream_t stream = nullptr;
    bool own_stream = false;
    mipt_hip_batch_config cfg{};
    std::vector<mipt_hip_traj_spec> specs;
    std::vector<TrajDesc> desc;
    std::vector<int> gen_list;  // trajectories that own their gates
    DevArena a;
    int *d_traj_list = nullptr;  // all trajectory indices (grouped order)
    std::vector<int> order;      // grouped order of trajectories
    std::size_t states_elems = 0, gates_count = 0, layer_items = 0, unif_items = 0;
    std::vector<std::uint32_t> h_flags;  // host mirror (for overflow planning)
    bool flags_on_host = false;
    std::vector<char> initialized;  // per trajectory: state already set by caller

    // Tiled groups alternate between two streams so that one group's small
    // finalize kernels and launch gaps overlap the other group's sweeps.
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipStream_t ls = nullptr;  // stream of the group being issued
    DevArena al{};             // arena view of that group (its own partials buffer)
    double *partials2 = nullptr;
    std::size_t partial2_cap = 0;
    // per stream lane: each layer's trajectory order for layer fusion
    int *fuse_lists[2] = {nullptr, nullptr};
    std::size_t fuse_cap[2] = {0, 0};

    // Sampled HIP-event timing of the gate sweeps (pass kernels).
    int prof_every = 0;
    int streams = 2;  // tiled groups in flight (1: every group on the batch stream)
    long prof_counter = 0;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, double>> ev_used;  // (event pair index, algorithmic bytes)
    long gate_launches = 0;
    double gate_bytes = 0.0;

    // the batch's own entropy workspace (RDM buffers, eigensolver panels, the
    // rocSOLVER handle), created with the first readout, freed with the batch
    Workspace *ews = nullptr;
    Workspace &ew()
    {
        if (!ews) ews = workspace_create();
        return *ews;
    }

    ~Batch()
    {
        drop_graph();
        release();
        workspace_destroy(ews);
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        if (partials2) (void)hipFree(partials2);
        for (int *fl : fuse_lists)
            if (fl) (void)hipFree(fl);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (side) (void)hipStreamDestroy(side);
    }

    int timed_launch_begin(double bytes, int &slot)
    {
        slot = -1;
        ++gate_launches;
        gate_bytes += bytes;
        if (prof_every <= 0 || ls != stream || (prof_counter++ % prof_every) != 0) return OK;
        const int pair = static_cast<int>(ev_used.size());
        while (ev_pool.size() < static_cast<std::size_t>(2 * pair + 2))
        {
            hipEvent_t e;
            MIPT_HIP_CHECK(hipEventCreate(&e));
            ev_pool.push_back(e);
        }
        MIPT_HIP_CHECK(hipEventRecord(ev_pool[2 * pair], stream));
        ev_used.push_back({pair, bytes});
        slot = pair;
        return OK;
    }

    int timed_launch_end(int slot)
    {
        if (slot < 0) return OK;
        MIPT_HIP_CHECK(hipEventRecord(ev_pool[2 * slot + 1], stream));
        return OK;
    }

    void release()
    {
        void *ptrs[] = {a.states, a.gates, a.desc, a.flags, a.outcomes, a.flagged, a.uniforms,
                        a.p1, a.coll, a.surprisal, a.status, a.init, a.partials, d_traj_list};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        a = DevArena{};
        d_traj_list = nullptr;
        if (own_stream && stream) (void)hipStreamDestroy(stream);
        own_stream = false;
    }

    // Tile variants: 14 -> 2^14-amplitude tiles, 256 threads x 64 registers
    // (one workgroup per CU); 13 -> 2^13 tiles, 256 x 32 (two per CU), or
    // 128 x 64 with threads = 128; 12 -> 64 x 64.
    std::map<int, Config> tuned;  // per qubit count, picked by autotune()
    // graph capture of the full evolution (set_graph): captured on the
    // second full evolve after the draws last changed, replayed afterwards
    bool graph_on = false;
    int graph_warm = 0;  // full eager evolves since the draws/config changed
    int graph_fuse = -1; // layer fusion setting the graph was captured with
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;

    void drop_graph()
    {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_warm = 0;
    }

    Config config_for(int nq) const
    {
        if (cfg.tile_bits == 0)
        {
            const auto it = tuned.find(nq);
            if (it != tuned.end()) return it->second;
        }
        Config cf{5, 256, 5};
        if (cfg.tile_bits == 14) cf = Config{6, 256, 6};
        if (cfg.tile_bits == 13) cf = cfg.threads == 128 ? Config{6, 128, 5} : Config{5, 256, 5};
        if (cfg.tile_bits == 12) cf = Config{6, 64, 4};
        // 32-register tiles default to matrix-core bonds (the faster form
        // measured on MI355X); 64-register tiles are built for the VALU only
        if (cf.rb == 5) cf.mx = cfg.bond_arith > 0 ? cfg.bond_arith - 1 : 2;
        cf.plan = cfg.plan;
        return cf;
    }

    int setup(const mipt_hip_traj_spec *sp, int count)
    {
        drop_graph();
        tuned.clear();
        void *ptrs[] = {a.states, a.gates, a.desc, a.flags, a.outcomes, a.flagged, a.uniforms,
                        a.p1, a.coll, a.surprisal, a.status, a.init, a.partials, d_traj_list};
        for (void *p : ptrs)
            if (p) MIPT_HIP_CHECK(hipFree(p));
        a = DevArena{};
        d_traj_list = nullptr;
        specs.assign(sp, sp + count);
        desc.resize(count);
        gen_list.clear();
        std::size_t so = 0, go = 0, lo = 0, uo = 0;
        for (int i = 0; i < count; ++i)
        {
            const auto &s = specs[i];
            // up to 2^27 amplitudes per state (1 GiB): tile byte offsets stay below 2^31
            const bool grid = s.grid_x > 0;
            if (s.L < 2 || s.L > 26 || (!grid && (s.L % 2) != 0) || s.layers < 0 || s.gate_family != 1 ||
                (s.probe_site >= s.L) || s.grid_x < 0 || !probe_spec_ok(s) ||
                (grid && (s.grid_x < 2 || (s.L % s.grid_x) != 0 || s.L / s.grid_x < 2)) ||
                (s.shared_gates >= 0 && (s.shared_gates >= count || s.shared_gates == i)))
            {
                set_error("batch_setup: unsupported spec %d (L=%d layers=%d family=%d)", i, s.L, s.layers, s.gate_family);
                return E_ARG;
            }
            TrajDesc &d = desc[i];
            d.L = s.L;
            d.probes = s.probe_count > 0 ? s.probe_count : (s.probe_site >= 0 ? 1 : 0);
            for (int j = 0; j < 4; ++j) d.psite[j] = -1;
            if (s.probe_count > 0)
                for (int j = 0; j < s.probe_count; ++j) d.psite[j] = s.probe_sites[j];
            else if (s.probe_site >= 0)
                d.psite[0] = s.probe_site;
            d.nq = s.L + d.probes;
            d.attach = (s.probe_count > 0 && s.attach_layer >= 0) ? s.attach_layer : s.p_zero_layers;
            d.stagger = s.probe_count > 0 ? s.attach_stagger : 0;
            d.layers = s.layers;
            d.closed = s.closed ? 1 : 0;
            d.gx = s.grid_x;
            d.probe_site = d.psite[0];
            d.p_zero_layers = s.p_zero_layers;
            d.p = s.p;
            d.seed = s.seed;
            d.state_off = so;
            so += std::size_t{1} << d.nq;
            d.layer_off = static_cast<std::uint32_t>(lo);
            lo += s.layers;
            d.unif_off = static_cast<std::uint32_t>(uo);
            uo += static_cast<std::size_t>(s.layers) * s.L;
            if (s.shared_gates < 0)
            {
                d.gate_off = go;
                std::size_t nb = 0;
                for (int t = 0; t < s.layers; ++t) nb += geo_bond_count(s.L, d.gx, t, d.closed);
                go += nb;
                gen_list.push_back(i);
            }
        }
        for (int i = 0; i < count; ++i)
            if (specs[i].shared_gates >= 0)
            {
                const auto &o = specs[specs[i].shared_gates];
                if (o.shared_gates >= 0 || o.L != specs[i].L || o.layers < specs[i].layers || o.closed != specs[i].closed ||
                    o.grid_x != specs[i].grid_x)
                {
                    set_error("batch_setup: shared_gates of %d must name an owning trajectory of the same shape", i);
                    return E_ARG;
                }
                desc[i].gate_off = desc[specs[i].shared_gates].gate_off;
            }
        states_elems = so;
        gates_count = go;
        layer_items = lo;
        unif_items = uo;
        MIPT_HIP_CHECK(hipMalloc(&a.states, std::max<std::size_t>(so, 1) * sizeof(float2)));
        MIPT_HIP_CHECK(hipMalloc(&a.gates, std::max<std::size_t>(go, 1) * 16 * sizeof(float2)));
        MIPT_HIP_CHECK(hipMalloc(&a.desc, count * sizeof(TrajDesc)));
        MIPT_HIP_CHECK(hipMalloc(&a.flags, std::max<std::size_t>(lo, 1) * 4));
        MIPT_HIP_CHECK(hipMalloc(&a.outcomes, std::max<std::size_t>(lo, 1) * 4));
        MIPT_HIP_CHECK(hipMalloc(&a.flagged, std::max<std::size_t>(lo, 1) * 4));
        MIPT_HIP_CHECK(hipMalloc(&a.coll, std::max<std::size_t>(lo, 1) * sizeof(Collapse)));
        MIPT_HIP_CHECK(hipMalloc(&a.uniforms, std::max<std::size_t>(uo, 1) * 8));
        MIPT_HIP_CHECK(hipMalloc(&a.p1, std::max<std::size_t>(uo, 1) * 8));
        MIPT_HIP_CHECK(hipMalloc(&a.surprisal, count * 8));
        MIPT_HIP_CHECK(hipMalloc(&a.status, count * 4));
        MIPT_HIP_CHECK(hipMalloc(&a.init, count * 4));
        hipLaunchKernelGGL(fill_int_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, a.init, count, 1);
        MIPT_HIP_CHECK(hipGetLastError());
        MIPT_HIP_CHECK(hipMemcpyAsync(a.desc, desc.data(), count * sizeof(TrajDesc), hipMemcpyHostToDevice, stream));
        MIPT_HIP_CHECK(hipMemsetAsync(a.surprisal, 0, count * 8, stream));
        MIPT_HIP_CHECK(hipMemsetAsync(a.status, 0, count * 4, stream));
        MIPT_HIP_CHECK(hipMemsetAsync(a.p1, 0, std::max<std::size_t>(uo, 1) * 8, stream));
        MIPT_HIP_CHECK(hipMemsetAsync(a.outcomes, 0, std::max<std::size_t>(lo, 1) * 4, stream));
        MIPT_HIP_CHECK(hipMemsetAsync(a.flagged, 0, std::max<std::size_t>(lo, 1) * 4, stream));
        // grouped order: by qubit count, then index
        order.resize(count);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            const TrajDesc &a0 = desc[x], &b0 = desc[y];
            if (a0.nq != b0.nq) return a0.nq < b0.nq;
            if (a0.L != b0.L) return a0.L < b0.L;
            if (a0.closed != b0.closed) return a0.closed < b0.closed;
            if (a0.gx != b0.gx) return a0.gx < b0.gx;
            return a0.layers < b0.layers;
        });
        MIPT_HIP_CHECK(hipMalloc(&d_traj_list, count * sizeof(int)));
        MIPT_HIP_CHECK(hipMemcpyAsync(d_traj_list, order.data(), count * sizeof(int), hipMemcpyHostToDevice, stream));
        h_flags.assign(lo, 0);
        flags_on_host = false;
        initialized.assign(count, 0);
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        return OK;
    }

    int generate()
    {
        const int count = static_cast<int>(desc.size());
        if (count == 0) return OK;
        // gates
        std::vector<std::uint64_t> item_prefix(gen_list.size() + 1, 0);
        for (std::size_t i = 0; i < gen_list.size(); ++i)
        {
            const TrajDesc &d = desc[gen_list[i]];
            std::size_t nb = 0;
            for (int t = 0; t < d.layers; ++t) nb += geo_bond_count(d.L, d.gx, t, d.closed);
            item_prefix[i + 1] = item_prefix[i] + nb;
        }
        std::vector<std::uint64_t> layer_prefix(count + 1, 0);
        for (int i = 0; i < count; ++i) layer_prefix[i + 1] = layer_prefix[i] + desc[i].layers;
        std::uint64_t *d_ip = nullptr, *d_lp = nullptr;
        int *d_gl = nullptr;
        MIPT_HIP_CHECK(hipMalloc(&d_ip, item_prefix.size() * 8));
        MIPT_HIP_CHECK(hipMalloc(&d_lp, layer_prefix.size() * 8));
        MIPT_HIP_CHECK(hipMalloc(&d_gl, std::max<std::size_t>(gen_list.size(), 1) * 4));
        MIPT_HIP_CHECK(hipMemcpyAsync(d_ip, item_prefix.data(), item_prefix.size() * 8, hipMemcpyHostToDevice, stream));
        MIPT_HIP_CHECK(hipMemcpyAsync(d_lp, layer_prefix.data(), layer_prefix.size() * 8, hipMemcpyHostToDevice, stream));
        if (!gen_list.empty())
            MIPT_HIP_CHECK(hipMemcpyAsync(d_gl, gen_list.data(), gen_list.size() * 4, hipMemcpyHostToDevice, stream));
        const std::uint64_t ng = item_prefix.back();
        if (ng)
        {
            hipLaunchKernelGGL(gen_gates_kernel, dim3(static_cast<unsigned>((ng + 255) / 256)), dim3(256), 0, stream,
                               a.desc, d_gl, static_cast<int>(gen_list.size()), nullptr, a.gates, ng, d_ip);
            MIPT_HIP_CHECK(hipGetLastError());
        }
        const std::uint64_t nl = layer_prefix.back();
        if (nl)
        {
            hipLaunchKernelGGL(gen_flags_kernel, dim3(static_cast<unsigned>((nl + 255) / 256)), dim3(256), 0, stream,
                               a.desc, count, a.flags, a.uniforms, nl, d_lp);
            MIPT_HIP_CHECK(hipGetLastError());
        }
        MIPT_HIP_CHECK(hipMemcpyAsync(h_flags.data(), a.flags, layer_items * 4, hipMemcpyDeviceToHost, stream));
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        flags_on_host = true;
        MIPT_HIP_CHECK(hipFree(d_ip));
        MIPT_HIP_CHECK(hipFree(d_lp));
        MIPT_HIP_CHECK(hipFree(d_gl));
        return OK;
    }

    int set_layers(int traj, const double *g, const std::uint8_t *flags, const double *unif)
    {
        drop_graph();
        if (traj < 0 || traj >= static_cast<int>(desc.size()) || specs[traj].shared_gates >= 0)
        {
            set_error("set_layers: bad trajectory (or it shares another's gates)");
            return E_ARG;
        }
        const TrajDesc &d = desc[traj];
        std::size_t nb = 0;
        for (int t = 0; t < d.layers; ++t) nb += geo_bond_count(d.L, d.gx, t, d.closed);
        std::vector<float2> gf(nb * 16);
        for (std::size_t i = 0; i < nb * 16; ++i)
            gf[i] = make_float2(static_cast<float>(g[2 * i]), static_cast<float>(g[2 * i + 1]));
        std::vector<std::uint32_t> fm(d.layers, 0);
        for (int t = 0; t < d.layers; ++t)
            for (int s = 0; s < d.L; ++s)
                if (flags[static_cast<std::size_t>(t) * d.L + s]) fm[t] |= 1u << s;
        for (std::size_t i = 0; i < static_cast<std::size_t>(d.layers) * d.L; ++i)
            if (flags[i] && !(unif[i] >= 0.0 && unif[i] < 1.0))
            {
                set_error("set_layers: uniform out of [0,1)");
                return E_ARG;
            }
        MIPT_HIP_CHECK(hipMemcpyAsync(a.gates + d.gate_off * 16, gf.data(), gf.size() * sizeof(float2),
                                      hipMemcpyHostToDevice, stream));
        MIPT_HIP_CHECK(hipMemcpyAsync(a.flags + d.layer_off, fm.data(), fm.size() * 4, hipMemcpyHostToDevice, stream));
        MIPT_HIP_CHECK(hipMemcpyAsync(a.uniforms + d.unif_off, unif, static_cast<std::size_t>(d.layers) * d.L * 8,
                                      hipMemcpyHostToDevice, stream));
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        std::copy(fm.begin(), fm.end(), h_flags.begin() + d.layer_off);
        flags_on_host = true;
        return OK;
    }

    int set_state(int traj, const float *psi)
    {
        if (traj < 0 || traj >= static_cast<int>(desc.size()) || !psi) { set_error("set_state: bad arguments"); return E_ARG; }
        const TrajDesc &d = desc[traj];
        MIPT_HIP_CHECK(hipMemcpyAsync(a.states + d.state_off, psi, (std::size_t{1} << d.nq) * 8, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(1), 0, stream, a.init + traj, 1, 0);
        MIPT_HIP_CHECK(hipGetLastError());
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        initialized[traj] = 1;
        return OK;
    }

    int get_state(int traj, float *psi)
    {
        if (traj < 0 || traj >= static_cast<int>(desc.size()) || !psi) { set_error("get_state: bad arguments"); return E_ARG; }
        const TrajDesc &d = desc[traj];
        MIPT_HIP_CHECK(hipMemcpyAsync(psi, a.states + d.state_off, (std::size_t{1} << d.nq) * 8, hipMemcpyDeviceToHost, stream));
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        return OK;
    }

    int ensure_partials(std::size_t doubles, int lane)
    {
        double *&p = lane ? partials2 : a.partials;
        std::size_t &cap = lane ? partial2_cap : a.partial_cap;
        if (doubles <= cap) return OK;
        if (p) MIPT_HIP_CHECK(hipFree(p));
        p = nullptr;
        MIPT_HIP_CHECK(hipMalloc(&p, doubles * 8));
        cap = doubles;
        return OK;
    }

    template <int RB, int NT, int AUX, int MX, int ZP, int PAT = 0, int PAT2 = 0>
    int launch_pass(const PassPlan &pl, const PassPlan *next, const int *list, int ntraj, int layer)
    {
        PassArgs pa{al, list, ntraj, layer, pl, next ? *next : PassPlan{}};
        // half a tile, then the matrix-core coefficient table (of both sweeps when fused)
        const int slots = 2 * (pl.nstages + (next ? next->nstages : 0));
        const std::size_t lds = (std::size_t{1} << (pl.k - 1)) * sizeof(float2) +
                                (MX > 0 ? static_cast<std::size_t>(slots) * MX_SLOT_BYTES : 0);
        static bool attr_set = false;
        if (!attr_set)
        {
            MIPT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&pass_kernel<RB, NT, AUX, MX, ZP, PAT, PAT2>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               static_cast<int>((std::size_t{1} << 13) * sizeof(float2) +
                                                                2 * MX_MAX_SLOTS * MX_SLOT_BYTES)));
            attr_set = true;
        }
        const unsigned grid = static_cast<unsigned>(static_cast<long>(ntraj) * pl.ntiles);
        hipLaunchKernelGGL((pass_kernel<RB, NT, AUX, MX, ZP, PAT, PAT2>), dim3(grid), dim3(NT), lds, ls, pa,
                           static_cast<const int *>(list), static_cast<const TrajDesc *>(al.desc),
                           static_cast<const std::uint32_t *>(al.flags), static_cast<const Collapse *>(al.coll),
                           static_cast<const int *>(al.init), static_cast<const float2 *>(al.gates));
        MIPT_HIP_CHECK(hipGetLastError());
        return OK;
    }

    // Stage-free sweeps (AUX) have no bond arithmetic; gate sweeps of 2^13
    // tiles come in all three arithmetic mixes, the 64-register shapes on the
    // VALU (the matrix-core operands do not fit beside 64 amplitudes).
    // The stage holding a bond of qubit 0 is the plan's last (1 / 2: its
    // form), or none (0); any other plan is refused.
    int qubit0_form(const PassPlan &pl, int &zp)
    {
        zp = 0;
        for (int si = 0; si < pl.nstages; ++si)
            if (pl.st[si].zpair)
            {
                if (si != pl.nstages - 1)
                {
                    set_error("tile plan: the bond of qubit 0 is not in the last stage");
                    return E_UNSUPPORTED;
                }
                zp = pl.st[si].zpair;
            }
        return OK;
    }

    // The stage-structure code of plan pl (pat_*), or 0 when it has none.
    static int pattern_code(const PassPlan &pl)
    {
        if (pl.nstages < 1 || pl.nstages > 3) return 0;
        int c = pl.nstages;
        for (int si = 0; si < pl.nstages; ++si)
        {
            const StagePlan &sp = pl.st[si];
            if ((si > 0 && !sp.use_lds) || (sp.zpair && (si != pl.nstages - 1 || sp.bond[0] < 0))) return 0;
            c |= ((sp.use_lds ? 1 : 0) << (2 + 3 * si)) | ((sp.bond[0] >= 0 ? 1 : 0) << (3 + 3 * si)) |
                 ((sp.bond[1] >= 0 ? 1 : 0) << (4 + 3 * si));
        }
        return c;
    }

    // Launches the compiled-structure kernel (matrix-core bonds) for (zp,
    // pat, pat2) if it is one of kCompiledSweeps, else leaves found false.
    template <int I = 0>
    int launch_compiled(int zp, int pat, int pat2, const PassPlan &pl, const PassPlan *next, const int *list, int ntraj,
                        int layer, bool &found)
    {
        if constexpr (I < kNumCompiledSweeps)
        {
            constexpr CompiledSweep k = kCompiledSweeps[I];
            if (zp == k.zp && pat == k.pat && pat2 == k.pat2)
            {
                found = true;
                return launch_pass<5, 256, 0, 2, k.zp, k.pat, k.pat2>(pl, next, list, ntraj, layer);
            }
            return launch_compiled<I + 1>(zp, pat, pat2, pl, next, list, ntraj, layer, found);
        }
        return OK;
    }

    // next (with pl.fuse): the next layer's first sweep, run in the same
    // launch for the trajectories that fuse this layer into the next.
    int run_pass(const Config &cf, const PassPlan &pl, const int *list, int ntraj, int layer,
                 const PassPlan *next = nullptr)
    {
        const bool aux = pl.nstages == 0;
        if (cf.nt == 256 && cf.rb == 5)
        {
            if (aux) return launch_pass<5, 256, 1, 0, 0>(pl, nullptr, list, ntraj, layer);
            int zp = 0;
            MIPT_HIP_TRY(qubit0_form(pl, zp));
            if (pl.fuse)
            {
                int zn = 0;
                MIPT_HIP_TRY(qubit0_form(*next, zn));
                if (zp != 0 || cf.mx == 1)
                {
                    set_error("layer fusion needs a first sweep without the bond of qubit 0");
                    return E_UNSUPPORTED;
                }
                if (cf.mx == 0)
                    return zn == 0   ? launch_pass<5, 256, 0, 0, 4>(pl, next, list, ntraj, layer)
                           : zn == 1 ? launch_pass<5, 256, 0, 0, 5>(pl, next, list, ntraj, layer)
                                     : launch_pass<5, 256, 0, 0, 6>(pl, next, list, ntraj, layer);
                bool found = false;
                MIPT_HIP_TRY(launch_compiled(4 + zn, pattern_code(pl), pattern_code(*next), pl, next, list, ntraj, layer,
                                             found));
                if (found) return OK;
                return zn == 0   ? launch_pass<5, 256, 0, 2, 4>(pl, next, list, ntraj, layer)
                       : zn == 1 ? launch_pass<5, 256, 0, 2, 5>(pl, next, list, ntraj, layer)
                                 : launch_pass<5, 256, 0, 2, 6>(pl, next, list, ntraj, layer);
            }
            if (cf.mx == 0)
                return zp == 0   ? launch_pass<5, 256, 0, 0, 0>(pl, nullptr, list, ntraj, layer)
                       : zp == 1 ? launch_pass<5, 256, 0, 0, 1>(pl, nullptr, list, ntraj, layer)
                                 : launch_pass<5, 256, 0, 0, 2>(pl, nullptr, list, ntraj, layer);
            if (cf.mx == 2)
            {
                bool found = false;
                MIPT_HIP_TRY(launch_compiled(zp, pattern_code(pl), 0, pl, nullptr, list, ntraj, layer, found));
                if (found) return OK;
            }
            if (cf.mx == 2)
                return zp == 0   ? launch_pass<5, 256, 0, 2, 0>(pl, nullptr, list, ntraj, layer)
                       : zp == 1 ? launch_pass<5, 256, 0, 2, 1>(pl, nullptr, list, ntraj, layer)
                                 : launch_pass<5, 256, 0, 2, 2>(pl, nullptr, list, ntraj, layer);
            return zp == 0   ? launch_pass<5, 256, 0, 1, 0>(pl, nullptr, list, ntraj, layer)
                   : zp == 1 ? launch_pass<5, 256, 0, 1, 1>(pl, nullptr, list, ntraj, layer)
                             : launch_pass<5, 256, 0, 1, 2>(pl, nullptr, list, ntraj, layer);
        }
        if (pl.fuse)
        {
            set_error("layer fusion needs 2^13-amplitude tiles of 256 threads");
            return E_UNSUPPORTED;
        }
        if (cf.nt == 256)
            return aux ? launch_pass<6, 256, 1, 0, 0>(pl, nullptr, list, ntraj, layer)
                       : launch_pass<6, 256, 0, 0, 3>(pl, nullptr, list, ntraj, layer);
        if (cf.nt == 128)
            return aux ? launch_pass<6, 128, 1, 0, 0>(pl, nullptr, list, ntraj, layer)
                       : launch_pass<6, 128, 0, 0, 3>(pl, nullptr, list, ntraj, layer);
        return aux ? launch_pass<6, 64, 1, 0, 0>(pl, nullptr, list, ntraj, layer)
                   : launch_pass<6, 64, 0, 0, 3>(pl, nullptr, list, ntraj, layer);
    }

    // Tiled evolution of a group of same-size trajectories (nq > 14).
    int evolve_group_tiled(const std::vector<int> &grp, int grp_pos, int t0, int t1, int lane)
    {
        const TrajDesc &d0 = desc[grp[0]];
        const Config cf = config_for(d0.nq);
        const int P = period_layers(d0.gx);
        LayerPlan lp[4];
        for (int j = 0; j < P; ++j) MIPT_HIP_TRY(build_layer_plan(d0.L, d0.nq, d0.closed, d0.gx, j, cf, lp[j]));
        const int ntraj = static_cast<int>(grp.size());
        const int *list = d_traj_list + grp_pos;
        MIPT_HIP_TRY(ensure_partials(static_cast<std::size_t>(ntraj) * (std::size_t{1} << (d0.nq - cf.k())) * HSTRIDE,
                                     lane));
        ls = lane ? side : stream;
        al = a;
        al.partials = lane ? partials2 : a.partials;
        // maximum measured count per layer over the group (host flags)
        bool attach_any = false;
        for (int i : grp)
            attach_any |= desc[i].probes > 0 && (desc[i].attach > 0 || (desc[i].probes > 1 && desc[i].stagger > 0));
        // layer fusion (2^13 tiles of 256 threads, no reference attached
        // mid-evolution): the layers' sweeps alternate their window order so
        // that consecutive layers meet in one window; a trajectory that
        // measures at most FUSE_MAX_MEASURED sites in layer t, none bonded in
        // layer t+1's first sweep (fuses()), runs that sweep in the same
        // launch as layer t's last (one pass over its state fewer), and its
        // collapse of layer t at layer t+1's second sweep; the others run
        // the two sweeps as usual.  Each launch gets a compacted list
        // (fuse_order_kernel), so no workgroup idles.
        bool fuse_env = false;
        MIPT_HIP_TRY(fuse_layers(fuse_env));
        // (the order kernel keeps a byte per trajectory of the group in LDS:
        // groups beyond 64 KiB of them, possible only with a raised
        // MIPT_HIP_GROUP_MIB, run unfused)
        const bool fuse_on = fuse_env && cf.rb == 5 && cf.nt == 256 && cf.mx != 1 && !attach_any && t1 - t0 > 1 &&
                             ntraj <= 65536;
        FuseRule rule{};
        rule.period = P;
        rule.t1 = t1;
        PassPlan nxp[4]{};
        for (int j = 0; j < P && fuse_on; ++j)
        {
            // layer parity j: its last sweep and the next layer's first share a
            // window, and neither holds the bond of qubit 0 in a way the fused
            // kernel cannot take
            const LayerPlan &LJ = lp[j];
            const LayerPlan &LN = lp[(j + 1) % P];
            const int np = static_cast<int>(LJ.passes.size()), nn = static_cast<int>(LN.passes.size());
            if (np < 2 || nn < 2) continue;
            const PassPlan &lastp = LJ.passes[(j & 1) ? 0 : np - 1];
            const PassPlan &nx = LN.passes[((j + 1) & 1) ? nn - 1 : 0];
            int z1 = 0;
            for (int si = 0; si < lastp.nstages; ++si) z1 |= lastp.st[si].zpair;
            rule.ok[j] = nx.k == lastp.k && nx.c == lastp.c && nx.s == lastp.s && nx.rest_mask == lastp.rest_mask &&
                         z1 == 0 && nx.nstages > 0 && nx.st[0].use_lds;
            rule.block[j] = nx.bonded;
            nxp[j] = nx;
        }
        int *qorder = nullptr;
        if (fuse_on)
        {
            const std::size_t need = static_cast<std::size_t>(ntraj) * (t1 - t0);
            if (need > fuse_cap[lane])
            {
                if (fuse_lists[lane]) MIPT_HIP_CHECK(hipFree(fuse_lists[lane]));
                fuse_lists[lane] = nullptr;
                MIPT_HIP_CHECK(hipMalloc(&fuse_lists[lane], need * sizeof(int)));
                fuse_cap[lane] = need;
            }
            qorder = fuse_lists[lane];
            hipLaunchKernelGGL(fuse_order_kernel, dim3(t1 - t0), dim3(256), static_cast<std::size_t>(ntraj), ls, list,
                               ntraj, static_cast<const TrajDesc *>(a.desc), static_cast<const std::uint32_t *>(a.flags),
                               t0, rule, qorder);
            MIPT_HIP_CHECK(hipGetLastError());
        }
        const double layer_bytes = 16.0 * static_cast<double>(std::size_t{1} << d0.nq);
        // trajectories that fused layer t-1 (they lead that layer's order and
        // already ran layer t's first sweep)
        int prev_nf = 0;
        for (int t = t0; t < t1; ++t)
        {
            const LayerPlan &LP = lp[t % P];
            if (attach_any && t > 0)
            {
                bool at = false;
                for (int i : grp)
                    for (int j = 0; j < desc[i].probes; ++j) at |= desc[i].attach + j * desc[i].stagger == t;
                if (at)
                {
                    if (t > t0) MIPT_HIP_TRY(close_layer(cf, lp[(t - 1) % P], list, ntraj, t));
                    hipLaunchKernelGGL(attach_kernel, dim3(ntraj), dim3(SMALL_THREADS), 0, ls, al, list, t);
                    MIPT_HIP_CHECK(hipGetLastError());
                }
            }
            int maxm = 0;
            for (int i : grp) maxm = std::max(maxm, __builtin_popcount(h_flags[desc[i].layer_off + t]));
            const int chunks = (maxm + MAX_HIST_BITS - 1) / MAX_HIST_BITS;
            const int np = static_cast<int>(LP.passes.size());
            int order[8];
            for (int i = 0; i < np; ++i) order[i] = fuse_on && (t & 1) ? np - 1 - i : i;
            const PassPlan &lastp = LP.passes[order[np - 1]];
            // this layer's order (fusing trajectories first, the others from
            // position nf), the same count the device derives from its flags
            const int *ord = fuse_on ? qorder + static_cast<std::size_t>(t - t0) * ntraj : list;
            int nf = 0;
            if (fuse_on)
                for (int i : grp) nf += fuses(rule, h_flags[desc[i].layer_off + t], t) ? 1 : 0;
            auto launch = [&](PassPlan pl, const int *lst, int cnt, double passes, const PassPlan *nxt) {
                if (cnt == 0) return static_cast<int>(OK);
                // algorithmic bytes: one read + one write of the states per
                // layer, shared among the layer's sweeps, for the sweeps run
                int slot;
                MIPT_HIP_TRY(timed_launch_begin(layer_bytes * cnt * passes / np, slot));
                MIPT_HIP_TRY(run_pass(cf, pl, lst, cnt, t, nxt));
                MIPT_HIP_TRY(timed_launch_end(slot));
                return static_cast<int>(OK);
            };
            for (int i = 0; i < np; ++i)
            {
                PassPlan pl = LP.passes[order[i]];
                pl.collapse = (i == 0) ? 1 : 0;
                pl.init = (i == 0 && t == 0) ? 1 : 0;
                pl.hist = (i + 1 == np) ? 1 : 0;  // chunk 0 always: it also records m == 0
                pl.chunk = 0;
                pl.cond = 0;
                pl.store = 1;
                // the collapse sweep skips zero-branch tiles when a later sweep
                // of the layer re-applies the branch on the sites it leaves alone
                pl.zskip = (i == 0 && np > 1) ? 1 : 0;
                pl.pmask = (i == 1) ? (((std::uint64_t{1} << d0.nq) - 1) & ~LP.passes[order[0]].bonded) : 0;
                pl.fuse = 0;
                pl.pend = (i == 1) ? 1 : 0;
                pl.pbase = 0;
                if (i == 0)
                {
                    // those that fused layer t-1 ran this sweep already
                    const int *others = prev_nf > 0 ? qorder + static_cast<std::size_t>(t - 1 - t0) * ntraj + prev_nf : ord;
                    MIPT_HIP_TRY(launch(pl, others, ntraj - prev_nf, 1.0, nullptr));
                }
                else if (i + 1 == np && nf > 0)
                {
                    PassPlan pr = pl;
                    pr.pbase = nf;
                    MIPT_HIP_TRY(launch(pr, ord + nf, ntraj - nf, 1.0, nullptr));
                    PassPlan pf = pl;
                    pf.fuse = 1;
                    MIPT_HIP_TRY(launch(pf, ord, nf, 2.0, &nxp[t % P]));
                }
                else MIPT_HIP_TRY(launch(pl, ord, ntraj, 1.0, nullptr));
            }
            // the histogram partials of the layer's last sweep sit at the
            // trajectories' positions in the layer's order
            const Geom geo{lastp.rest_mask, lastp.ntiles, lastp.k, lastp.c, lastp.s};
            FinalizeArgs fa{al, ord, t, 0, geo, 0};
            hipLaunchKernelGGL(finalize_kernel, dim3(ntraj), dim3(256), 0, ls, fa);
            MIPT_HIP_CHECK(hipGetLastError());
            // more than MAX_HIST_BITS measured sites: never a fusing trajectory
            for (int ch = 1; ch < chunks; ++ch)
            {
                PassPlan pl = lastp;
                pl.nstages = 0;
                pl.collapse = 0;
                pl.init = 0;
                pl.hist = 1;
                pl.chunk = ch;
                pl.cond = 1;
                pl.store = 0;
                pl.fuse = 0;
                pl.pend = 0;
                pl.pbase = nf;
                MIPT_HIP_TRY(run_pass(cf, pl, ord + nf, ntraj - nf, t));
                FinalizeArgs fc{al, ord + nf, t, ch, geo, nf};
                hipLaunchKernelGGL(finalize_kernel, dim3(ntraj - nf), dim3(256), 0, ls, fc);
                MIPT_HIP_CHECK(hipGetLastError());
            }
            prev_nf = nf;
        }
        // apply the last layer's pending collapse
        if (t1 > t0) MIPT_HIP_TRY(close_layer(cf, lp[(t1 - 1) % P], list, ntraj, t1));
        return OK;
    }

    // Applies layer t-1's pending collapse with a stage-free sweep and marks
    // it applied, so that the state is the post-measurement state before
    // layer t (the end of an evolution range, or a reference attach at t).
    int close_layer(const Config &cf, const LayerPlan &prev, const int *list, int ntraj, int t)
    {
        PassPlan pl = prev.passes[0];
        pl.nstages = 0;
        pl.collapse = 1;
        pl.init = 0;
        pl.hist = 0;
        pl.store = 1;
        MIPT_HIP_TRY(run_pass(cf, pl, list, ntraj, t));
        hipLaunchKernelGGL(reset_collapse_kernel, dim3((ntraj + 255) / 256), dim3(256), 0, ls, al, list, ntraj, t - 1);
        MIPT_HIP_CHECK(hipGetLastError());
        return OK;
    }

    // t1 < 0: every trajectory through its own depth.
    // Full evolutions may run as a replayed graph (see set_graph); the flags
    // and draws the graph was captured with are those on the host now, which
    // only set_layers()/setup() change (generate() redraws the same streams).
    int evolve(int t0, int t1_req)
    {
        if (!flags_on_host)
        {
            set_error("evolve: call generate() or set_layers() first");
            return E_ARG;
        }
        if (graph_on && t0 == 0 && t1_req < 0)
        {
            // a change of the layer-fusion setting changes the launch sequence
            // (and the buffers it needs): warm up and capture again
            bool fz = false;
            MIPT_HIP_TRY(fuse_layers(fz));
            if (static_cast<int>(fz) != graph_fuse)
            {
                drop_graph();
                graph_fuse = static_cast<int>(fz);
            }
            if (graph_exec)
            {
                MIPT_HIP_CHECK(hipGraphLaunch(graph_exec, stream));
                for (auto &x : initialized) x = 1;
                return OK;
            }
            if (graph_warm > 0)
            {
                // every buffer exists after an eager run: capture one
                const int saved_prof = prof_every;
                prof_every = 0;
                MIPT_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
                const int st = evolve_body(t0, t1_req);
                hipGraph_t gph = nullptr;
                const hipError_t ec = hipStreamEndCapture(stream, &gph);
                prof_every = saved_prof;
                if (st != OK)
                {
                    if (gph) (void)hipGraphDestroy(gph);
                    return st;
                }
                MIPT_HIP_CHECK(ec);
                graph = gph;
                MIPT_HIP_CHECK(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
                MIPT_HIP_CHECK(hipGraphLaunch(graph_exec, stream));
                return OK;
            }
            ++graph_warm;
        }
        return evolve_body(t0, t1_req);
    }

    int evolve_body(int t0, int t1_req)
    {
        const int count = static_cast<int>(order.size());
        if (!side)
        {
            MIPT_HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
            MIPT_HIP_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
            MIPT_HIP_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        }
        MIPT_HIP_CHECK(hipEventRecord(ev_fork, stream));
        MIPT_HIP_CHECK(hipStreamWaitEvent(side, ev_fork, 0));
        // tiled groups of every shape, dealt to the in-flight streams below
        struct TiledTask
        {
            int q, qe, t1;
            double cost;  // amplitude-layers x sweeps per layer
        };
        std::vector<TiledTask> tasks;
        double small_cost = 0.0;
        int pos = 0;
        while (pos < count)
        {
            const int nq = desc[order[pos]].nq;
            int end = pos;
            while (end < count && desc[order[end]].nq == nq && desc[order[end]].L == desc[order[pos]].L &&
                   desc[order[end]].closed == desc[order[pos]].closed && desc[order[end]].gx == desc[order[pos]].gx &&
                   (t1_req >= 0 || desc[order[end]].layers == desc[order[pos]].layers))
                ++end;
            const int t1 = t1_req >= 0 ? t1_req : desc[order[pos]].layers;
            if (t1 <= t0)
            {
                pos = end;
                continue;
            }
            if (nq <= SMALL_MAX_QUBITS)
            {
                const std::size_t lds = (std::size_t{1} << nq) * sizeof(float2);
                static bool attr = false;
                if (!attr)
                {
                    MIPT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&small_evolve_kernel),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       static_cast<int>((std::size_t{1} << SMALL_MAX_QUBITS) * sizeof(float2))));
                    attr = true;
                }
                hipLaunchKernelGGL(small_evolve_kernel, dim3(end - pos), dim3(SMALL_THREADS), lds, stream, a,
                                   d_traj_list + pos, t0, t1, t0 == 0 ? 1 : 0);
                MIPT_HIP_CHECK(hipGetLastError());
                small_cost += static_cast<double>(end - pos) * static_cast<double>(std::size_t{1} << nq) * (t1 - t0);
            }
            else
            {
                // groups of trajectories of at most the per-stream budget
                const std::size_t bytes = (std::size_t{1} << nq) * 8;
                static const std::size_t budget = [] {
                    const char *e = std::getenv("MIPT_HIP_GROUP_MIB");  // per-stream group budget
                    const long v = e ? std::atol(e) : 0;
                    return static_cast<std::size_t>(v > 0 ? v : 4096) << 20;
                }();
                const int g = static_cast<int>(std::max<std::size_t>(1, budget / bytes));
                const TrajDesc &d0 = desc[order[pos]];
                LayerPlan lp0;
                const int sweeps = build_layer_plan(d0.L, d0.nq, d0.closed, d0.gx, 0, config_for(d0.nq), lp0) == OK
                                       ? static_cast<int>(lp0.passes.size())
                                       : 2;
                for (int q = pos; q < end; q += g)
                {
                    const int qe = std::min(end, q + g);
                    tasks.push_back({q, qe, t1, static_cast<double>(qe - q) * static_cast<double>(std::size_t{1} << nq) *
                                                    (t1 - t0) * sweeps});
                }
            }
            pos = end;
        }
        // a mixed-size batch is one queue: groups longest first, each to the
        // stream with the least work so far (the small kernels run on the
        // batch stream), so the streams finish together
        std::stable_sort(tasks.begin(), tasks.end(), [](const TiledTask &x, const TiledTask &y) { return x.cost > y.cost; });
        double load[2] = {small_cost, 0.0};
        for (const TiledTask &tk : tasks)
        {
            const int lane = (streams > 1 && load[1] < load[0]) ? 1 : 0;
            load[lane] += tk.cost;
            std::vector<int> grp(order.begin() + tk.q, order.begin() + tk.qe);
            MIPT_HIP_TRY(evolve_group_tiled(grp, tk.q, t0, tk.t1, lane));
        }
        ls = stream;
        MIPT_HIP_CHECK(hipEventRecord(ev_join, side));
        MIPT_HIP_CHECK(hipStreamWaitEvent(stream, ev_join, 0));
        if (t1_req < 0 || t1_req > t0)
        {
            hipLaunchKernelGGL(fill_int_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, a.init, count, 0);
            MIPT_HIP_CHECK(hipGetLastError());
            for (auto &x : initialized) x = 1;
        }
        return OK;
    }

    // Re-arm in-kernel initial states and clear per-run records (a new run of
    // the same trajectories).
    // Autotuner: times `probe_layers` layers of this batch (its own draws and
    // shapes) under each tile configuration and keeps the fastest for the
    // tile-swept qubit counts; the batch is re-armed before and after, so the
    // draws stay and the next evolve starts from the initial states.
    int autotune(int probe_layers, int *best_index, int *best_plan, double *ms_out)
    {
        if (!flags_on_host)
        {
            set_error("autotune: call generate() or set_layers() first");
            return E_ARG;
        }
        if (cfg.tile_bits != 0)
        {
            set_error("autotune: the batch was created with a fixed tile configuration");
            return E_ARG;
        }
        int depth = probe_layers;
        std::vector<int> nqs;
        for (const TrajDesc &d : desc)
            if (d.nq > SMALL_MAX_QUBITS)
            {
                depth = std::min(depth, d.layers);
                if (std::find(nqs.begin(), nqs.end(), d.nq) == nqs.end()) nqs.push_back(d.nq);
            }
        if (best_index) *best_index = -1;
        if (best_plan) *best_plan = 0;
        if (ms_out)
            for (int c = 0; c < AUTOTUNE_CANDIDATES + AUTOTUNE_PLANS - 1; ++c) ms_out[c] = 0.0;
        if (nqs.empty() || depth < 1) return OK;  // nothing tile-swept
        const Config cand[AUTOTUNE_CANDIDATES] = {Config{5, 256, 5, 0}, Config{5, 256, 5, 1}, Config{5, 256, 5, 2},
                                                  Config{6, 256, 6, 0}, Config{6, 128, 5, 0}, Config{6, 64, 4, 0}};
        drop_graph();
        hipEvent_t e0, e1;
        MIPT_HIP_CHECK(hipEventCreate(&e0));
        MIPT_HIP_CHECK(hipEventCreate(&e1));
        int st = OK;
        // time one configuration on this batch's own draws (second of two runs);
        // < 0 when some layer plan does not exist for it
        auto probe = [&](const Config &cf, double &ms_best) -> int {
            for (int nq : nqs) tuned[nq] = cf;
            for (const TrajDesc &d : desc)
                if (d.nq > SMALL_MAX_QUBITS)
                    for (int j = 0; j < period_layers(d.gx); ++j)
                    {
                        LayerPlan lp;
                        if (build_layer_plan(d.L, d.nq, d.closed, d.gx, j, cf, lp) != OK) return -1;
                    }
            for (int rep = 0; rep < 2; ++rep)  // the first run warms buffers and code
            {
                MIPT_HIP_TRY(reset());
                MIPT_HIP_CHECK(hipEventRecord(e0, stream));
                MIPT_HIP_TRY(evolve_body(0, depth));
                MIPT_HIP_CHECK(hipEventRecord(e1, stream));
                MIPT_HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.f;
                MIPT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                ms_best = ms;
            }
            return OK;
        };
        // tile shapes and bond arithmetic on the default layer plans ...
        int best = -1;
        double best_ms = 0.0;
        for (int c = 0; c < AUTOTUNE_CANDIDATES && st == OK; ++c)
        {
            double ms = 0.0;
            const int r = probe(cand[c], ms);
            if (r < 0) continue;
            st = r;
            if (st != OK) break;
            if (ms_out) ms_out[c] = ms;
            if (best < 0 || ms < best_ms)
            {
                best = c;
                best_ms = ms;
            }
        }
        // ... then the qubit-block assignment (layer plan variants) of the winner
        int plan = 0;
        for (int v = 1; v < AUTOTUNE_PLANS && st == OK && best >= 0; ++v)
        {
            Config cf = cand[best];
            cf.plan = v;
            double ms = 0.0;
            const int r = probe(cf, ms);
            if (r < 0) continue;
            st = r;
            if (st != OK) break;
            if (ms_out) ms_out[AUTOTUNE_CANDIDATES + v - 1] = ms;
            if (ms < best_ms)
            {
                plan = v;
                best_ms = ms;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (st != OK) return st;
        Config win = cand[best < 0 ? 0 : best];
        win.plan = plan;
        for (int nq : nqs) tuned[nq] = win;
        if (best_index) *best_index = best;
        if (best_plan) *best_plan = plan;
        return reset();
    }

    int reset()
    {
        const int count = static_cast<int>(desc.size());
        if (!count) return OK;
        hipLaunchKernelGGL(fill_int_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, a.init, count, 1);
        MIPT_HIP_CHECK(hipGetLastError());
        MIPT_HIP_CHECK(hipMemsetAsync(a.surprisal, 0, count * 8, stream));
        MIPT_HIP_CHECK(hipMemsetAsync(a.status, 0, count * 4, stream));
        std::fill(initialized.begin(), initialized.end(), 0);
        return OK;
    }

    int export_layers(int traj, double *g, std::uint8_t *flags, double *unif)
    {
        if (traj < 0 || traj >= static_cast<int>(desc.size())) { set_error("export_layers: bad trajectory"); return E_ARG; }
        const TrajDesc &d = desc[traj];
        std::size_t nb = 0;
        for (int t = 0; t < d.layers; ++t) nb += geo_bond_count(d.L, d.gx, t, d.closed);
        if (g)
        {
            std::vector<float2> gf(nb * 16);
            MIPT_HIP_CHECK(hipMemcpyAsync(gf.data(), a.gates + d.gate_off * 16, gf.size() * sizeof(float2),
                                          hipMemcpyDeviceToHost, stream));
            MIPT_HIP_CHECK(hipStreamSynchronize(stream));
            for (std::size_t i = 0; i < gf.size(); ++i)
            {
                g[2 * i] = gf[i].x;
                g[2 * i + 1] = gf[i].y;
            }
        }
        std::vector<std::uint32_t> fm(d.layers);
        MIPT_HIP_CHECK(hipMemcpyAsync(fm.data(), a.flags + d.layer_off, d.layers * 4, hipMemcpyDeviceToHost, stream));
        if (unif)
            MIPT_HIP_CHECK(hipMemcpyAsync(unif, a.uniforms + d.unif_off, static_cast<std::size_t>(d.layers) * d.L * 8,
                                          hipMemcpyDeviceToHost, stream));
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        if (flags)
            for (int t = 0; t < d.layers; ++t)
                for (int q = 0; q < d.L; ++q) flags[static_cast<std::size_t>(t) * d.L + q] = (fm[t] >> q) & 1u;
        return OK;
    }

    int pass_stats(double *avg_ms, int *timed, double *bytes_avg, long *launches, double *bytes_total)
    {
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        double ms_sum = 0.0, b_sum = 0.0;
        for (const auto &u : ev_used)
        {
            float ms = 0.f;
            MIPT_HIP_CHECK(hipEventElapsedTime(&ms, ev_pool[2 * u.first], ev_pool[2 * u.first + 1]));
            ms_sum += ms;
            b_sum += u.second;
        }
        const int n = static_cast<int>(ev_used.size());
        if (avg_ms) *avg_ms = n ? ms_sum / n : 0.0;
        if (timed) *timed = n;
        if (bytes_avg) *bytes_avg = n ? b_sum / n : 0.0;
        if (launches) *launches = gate_launches;
        if (bytes_total) *bytes_total = gate_bytes;
        return OK;
    }

    int outcomes(int traj, int *out, std::uint8_t *flagged, double *p1, double *surprisal)
    {
        if (traj < 0 || traj >= static_cast<int>(desc.size())) { set_error("outcomes: bad trajectory"); return E_ARG; }
        const TrajDesc &d = desc[traj];
        std::vector<std::uint32_t> o(d.layers), f(d.layers), fl(d.layers);
        MIPT_HIP_CHECK(hipMemcpyAsync(o.data(), a.outcomes + d.layer_off, d.layers * 4, hipMemcpyDeviceToHost, stream));
        MIPT_HIP_CHECK(hipMemcpyAsync(f.data(), a.flagged + d.layer_off, d.layers * 4, hipMemcpyDeviceToHost, stream));
        MIPT_HIP_CHECK(hipMemcpyAsync(fl.data(), a.flags + d.layer_off, d.layers * 4, hipMemcpyDeviceToHost, stream));
        std::vector<double> pp(static_cast<std::size_t>(d.layers) * d.L);
        if (!pp.empty())
            MIPT_HIP_CHECK(hipMemcpyAsync(pp.data(), a.p1 + d.unif_off, pp.size() * 8, hipMemcpyDeviceToHost, stream));
        double sp = 0;
        int st = 0;
        MIPT_HIP_CHECK(hipMemcpyAsync(&sp, a.surprisal + traj, 8, hipMemcpyDeviceToHost, stream));
        MIPT_HIP_CHECK(hipMemcpyAsync(&st, a.status + traj, 4, hipMemcpyDeviceToHost, stream));
        MIPT_HIP_CHECK(hipStreamSynchronize(stream));
        for (int t = 0; t < d.layers; ++t)
            for (int s = 0; s < d.L; ++s)
            {
                const std::size_t i = static_cast<std::size_t>(t) * d.L + s;
                const bool meas = (fl[t] >> s) & 1u;
                if (out) out[i] = meas ? static_cast<int>((o[t] >> s) & 1u) : -1;
                if (flagged) flagged[i] = meas ? static_cast<std::uint8_t>((f[t] >> s) & 1u) : 0;
                if (p1) p1[i] = meas ? pp[i] : -1.0;
            }
        if (surprisal) *surprisal = sp;
        if (st != 0)
        {
            set_error("trajectory %d hit a zero-weight measurement branch", traj);
            return st;
        }
        return OK;
    }
};

} // namespace mipt_hip

using namespace mipt_hip;

struct mipt_hip_batch : Batch
{
};

extern "C" {

int mipt_hip_batch_create(const mipt_hip_batch_config *config, void *stream, mipt_hip_batch **out)
{
    if (!out) { set_error("batch_create: null out"); return E_ARG; }
    auto *b = new (std::nothrow) mipt_hip_batch();
    if (!b) return E_NOMEM;
    if (config) b->cfg = *config;
    if (b->cfg.tile_bits != 0 && (b->cfg.tile_bits < 12 || b->cfg.tile_bits > 14))
    {
        delete b;
        set_error("batch_create: tile_bits must be 0 or 12..14");
        return E_ARG;
    }
    if (b->cfg.plan < 0 || b->cfg.plan >= AUTOTUNE_PLANS)
    {
        delete b;
        set_error("batch_create: plan must be 0..%d", AUTOTUNE_PLANS - 1);
        return E_ARG;
    }
    if (b->cfg.bond_arith < 0 || b->cfg.bond_arith > 3 ||
        (b->cfg.bond_arith >= 2 && ((b->cfg.tile_bits != 0 && b->cfg.tile_bits != 13) || b->cfg.threads == 128)))
    {
        delete b;
        set_error("batch_create: bond_arith must be 0..3; 2 and 3 need 2^13 tiles of 256 threads");
        return E_ARG;
    }
    if (stream) b->stream = static_cast<hipStream_t>(stream);
    else
    {
        const hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (e != hipSuccess)
        {
            delete b;
            set_error("hipStreamCreate: %s", hipGetErrorString(e));
            return static_cast<int>(e);
        }
        b->own_stream = true;
    }
    *out = b;
    return OK;
}

int mipt_hip_batch_destroy(mipt_hip_batch *b)
{
    if (b) (void)hipStreamSynchronize(b->stream);
    delete b;
    return OK;
}

int mipt_hip_batch_setup(mipt_hip_batch *b, const mipt_hip_traj_spec *specs, int count)
{
    if (!b || !specs || count <= 0) { set_error("batch_setup: bad arguments"); return E_ARG; }
    return b->setup(specs, count);
}

int mipt_hip_batch_generate(mipt_hip_batch *b)
{
    if (!b) { set_error("batch_generate: null batch"); return E_ARG; }
    return b->generate();
}

int mipt_hip_batch_set_layers(mipt_hip_batch *b, int traj, const double *g, const uint8_t *f, const double *u)
{
    if (!b || !g || !f || !u) { set_error("batch_set_layers: null argument"); return E_ARG; }
    return b->set_layers(traj, g, f, u);
}

int mipt_hip_batch_set_state(mipt_hip_batch *b, int traj, const float *psi)
{
    if (!b) { set_error("batch_set_state: null batch"); return E_ARG; }
    return b->set_state(traj, psi);
}

int mipt_hip_batch_get_state(mipt_hip_batch *b, int traj, float *psi)
{
    if (!b) { set_error("batch_get_state: null batch"); return E_ARG; }
    return b->get_state(traj, psi);
}

void *mipt_hip_batch_state_ptr(mipt_hip_batch *b, int traj)
{
    if (!b || traj < 0 || traj >= static_cast<int>(b->desc.size())) return nullptr;
    return b->a.states + b->desc[traj].state_off;
}

int mipt_hip_batch_evolve(mipt_hip_batch *b)
{
    if (!b) { set_error("batch_evolve: null batch"); return E_ARG; }
    return b->evolve(0, -1);
}

int mipt_hip_batch_evolve_range(mipt_hip_batch *b, int t0, int t1)
{
    if (!b || t0 < 0 || t1 < t0) { set_error("batch_evolve_range: bad arguments"); return E_ARG; }
    for (const auto &d : b->desc)
        if (t1 > d.layers) { set_error("batch_evolve_range: layer beyond a trajectory's depth"); return E_ARG; }
    return b->evolve(t0, t1);
}

int mipt_hip_batch_autotune(mipt_hip_batch *b, int probe_layers, int *best_index, int *best_plan, double *ms)
{
    if (!b || probe_layers < 1) { set_error("batch_autotune: bad arguments"); return E_ARG; }
    return b->autotune(probe_layers, best_index, best_plan, ms);
}

int mipt_hip_batch_set_graph(mipt_hip_batch *b, int enable)
{
    if (!b) { set_error("batch_set_graph: null batch"); return E_ARG; }
    b->drop_graph();
    b->graph_on = enable != 0;
    return OK;
}

int mipt_hip_batch_synchronize(mipt_hip_batch *b)
{
    if (!b) { set_error("batch_synchronize: null batch"); return E_ARG; }
    MIPT_HIP_CHECK(hipStreamSynchronize(b->stream));
    return OK;
}

int mipt_hip_batch_records(mipt_hip_batch *b, int layer_begin, int layer_end, int8_t *host)
{
    if (!b || !host || b->desc.empty() || layer_begin < 0 || layer_end <= layer_begin)
    {
        set_error("batch_records: bad arguments");
        return E_ARG;
    }
    const int L = b->desc[0].L;
    for (const TrajDesc &d : b->desc)
        if (d.L != L || d.layers < layer_end)
        {
            set_error("batch_records: every trajectory needs L=%d and >= %d layers", L, layer_end);
            return E_ARG;
        }
    const int count = static_cast<int>(b->desc.size());
    std::vector<int> st(count);
    MIPT_HIP_CHECK(hipMemcpyAsync(st.data(), b->a.status, count * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    const std::size_t total = static_cast<std::size_t>(count) * (layer_end - layer_begin) * L;
    std::int8_t *dev = nullptr;
    MIPT_HIP_CHECK(hipMalloc(&dev, total));
    hipLaunchKernelGGL(records_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, b->stream, b->a,
                       layer_begin, layer_end - layer_begin, L, total, dev);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, dev, total, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(dev);
    if (e != hipSuccess)
    {
        set_error("batch_records: %s", hipGetErrorString(e));
        return static_cast<int>(e);
    }
    for (int i = 0; i < count; ++i)
        if (st[i] != 0)
        {
            set_error("trajectory %d hit a zero-weight measurement branch", i);
            return st[i];
        }
    return OK;
}

int mipt_hip_batch_rho3(mipt_hip_batch *b, int first, int count, const int *subsystems, int nsub, double *host)
{
    if (!b || !host || !subsystems || nsub < 1 || first < 0 || count < 1 ||
        first + count > static_cast<int>(b->desc.size()))
    {
        set_error("batch_rho3: bad arguments");
        return E_ARG;
    }
    for (int i = first; i < first + count; ++i)
    {
        const TrajDesc &d = b->desc[i];
        for (int k = 0; k < 3 * nsub; ++k)
            if (subsystems[k] < 0 || subsystems[k] >= d.L || d.nq < 4)
            {
                set_error("batch_rho3: subsystem qubit %d outside trajectory %d's %d sites", subsystems[k], i, d.L);
                return E_ARG;
            }
        for (int s = 0; s < nsub; ++s)
        {
            const int *q = subsystems + 3 * s;
            if (q[0] == q[1] || q[0] == q[2] || q[1] == q[2])
            {
                set_error("batch_rho3: subsystem %d repeats a qubit", s);
                return E_ARG;
            }
        }
    }
    int *dsub = nullptr;
    double *dev = nullptr;
    const std::size_t nout = static_cast<std::size_t>(count) * nsub * 128;
    MIPT_HIP_CHECK(hipMalloc(&dsub, 3 * nsub * sizeof(int)));
    hipError_t e = hipMalloc(&dev, nout * sizeof(double));
    if (e == hipSuccess) e = hipMemcpyAsync(dsub, subsystems, 3 * nsub * sizeof(int), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
    {
        hipLaunchKernelGGL(rho3_kernel, dim3(count * nsub), dim3(SMALL_THREADS), 0, b->stream, b->a, first, nsub, dsub,
                           dev);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(host, dev, nout * sizeof(double), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(dsub);
    if (dev) (void)hipFree(dev);
    if (e != hipSuccess)
    {
        set_error("batch_rho3: %s", hipGetErrorString(e));
        return static_cast<int>(e);
    }
    return OK;
}

int mipt_hip_batch_probe_bloch(mipt_hip_batch *b, double *host)
{
    if (!b || !host || b->desc.empty())
    {
        set_error("batch_probe_bloch: bad arguments");
        return E_ARG;
    }
    const int count = static_cast<int>(b->desc.size());
    double *dev = nullptr;
    MIPT_HIP_CHECK(hipMalloc(&dev, count * 3 * sizeof(double)));
    hipLaunchKernelGGL(probe_bloch_kernel, dim3(count), dim3(SMALL_THREADS), 0, b->stream, b->a, dev);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, dev, count * 3 * sizeof(double), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(dev);
    if (e != hipSuccess)
    {
        set_error("batch_probe_bloch: %s", hipGetErrorString(e));
        return static_cast<int>(e);
    }
    return OK;
}

int mipt_hip_batch_probe_rdm(mipt_hip_batch *b, double *host)
{
    if (!b || !host || b->desc.empty())
    {
        set_error("batch_probe_rdm: bad arguments");
        return E_ARG;
    }
    const int count = static_cast<int>(b->desc.size());
    const int k = b->desc[0].probes;
    for (const TrajDesc &d : b->desc)
        if (d.probes != k || k < 1)
        {
            set_error("batch_probe_rdm: every trajectory needs the same number (>= 1) of references");
            return E_ARG;
        }
    const int D = 1 << k, npair = D * (D + 1) / 2;
    double *dev = nullptr;
    MIPT_HIP_CHECK(hipMalloc(&dev, static_cast<std::size_t>(count) * npair * 2 * sizeof(double)));
    const int tiles = D < 4 ? 1 : D / 4;
    hipLaunchKernelGGL(probe_rdm_kernel, dim3(tiles * (tiles + 1) / 2, count), dim3(SMALL_THREADS), 0, b->stream, b->a,
                       D, dev);
    std::vector<double> tri(static_cast<std::size_t>(count) * npair * 2);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(tri.data(), dev, tri.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(dev);
    if (e != hipSuccess)
    {
        set_error("batch_probe_rdm: %s", hipGetErrorString(e));
        return static_cast<int>(e);
    }
    // Hermitian fill and division by the trace
    for (int t = 0; t < count; ++t)
    {
        const double *src = tri.data() + static_cast<std::size_t>(t) * npair * 2;
        double *dst = host + static_cast<std::size_t>(t) * D * D * 2;
        double tr = 0.0;
        for (int r = 0; r < D; ++r) tr += src[2 * (r * (r + 1) / 2 + r)];
        for (int r = 0; r < D; ++r)
            for (int c = 0; c <= r; ++c)
            {
                const int q = r * (r + 1) / 2 + c;
                const double re = src[2 * q] / tr, im = src[2 * q + 1] / tr;
                dst[2 * (r * D + c)] = re;
                dst[2 * (r * D + c) + 1] = r == c ? 0.0 : im;
                dst[2 * (c * D + r)] = re;
                dst[2 * (c * D + r) + 1] = r == c ? 0.0 : -im;
            }
    }
    return OK;
}

int mipt_hip_batch_outcomes(mipt_hip_batch *b, int traj, int *o, uint8_t *f, double *p1, double *s)
{
    if (!b) { set_error("batch_outcomes: null batch"); return E_ARG; }
    return b->outcomes(traj, o, f, p1, s);
}

void *mipt_hip_batch_stream(mipt_hip_batch *b) { return b ? static_cast<void *>(b->stream) : nullptr; }

int mipt_hip_batch_reset(mipt_hip_batch *b)
{
    if (!b) { set_error("batch_reset: null batch"); return E_ARG; }
    return b->reset();
}

int mipt_hip_batch_export_layers(mipt_hip_batch *b, int traj, double *g, uint8_t *f, double *u)
{
    if (!b) { set_error("batch_export_layers: null batch"); return E_ARG; }
    return b->export_layers(traj, g, f, u);
}

int mipt_hip_set_layer_fusion(int on)
{
    if (on < -1 || on > 1)
    {
        set_error("layer fusion must be 0 (off), 1 (on) or -1 (MIPT_HIP_FUSE / default), got %d", on);
        return E_ARG;
    }
    g_fuse.store(on);
    return OK;
}

int mipt_hip_compiled_sweeps(int *out, int out_len, int *n_out)
{
    if (!out || !n_out || out_len < 3 * kNumCompiledSweeps)
    {
        set_error("compiled_sweeps: need room for %d ints", 3 * kNumCompiledSweeps);
        return E_ARG;
    }
    for (int i = 0; i < kNumCompiledSweeps; ++i)
    {
        out[3 * i] = kCompiledSweeps[i].zp;
        out[3 * i + 1] = kCompiledSweeps[i].pat;
        out[3 * i + 2] = kCompiledSweeps[i].pat2;
    }
    *n_out = 3 * kNumCompiledSweeps;
    return OK;
}

int mipt_hip_layer_plan(int L, int nq, int closed, int grid_x, int layer, int tile_bits, int threads, int plan,
                        int *out, int out_len, int *n_out)
{
    if (!out || !n_out || L < 2 || nq < L || nq > 30 || layer < 0 || plan < 0 || plan > 3 || grid_x < 0 ||
        (grid_x > 0 && (L % grid_x) != 0))
    {
        set_error("layer_plan: bad arguments");
        return E_ARG;
    }
    Config cf;
    if (tile_bits == 14 && threads == 0) cf = Config{6, 256, 6};
    else if (tile_bits == 13 && (threads == 0 || threads == 256)) cf = Config{5, 256, 5};
    else if (tile_bits == 13 && threads == 128) cf = Config{6, 128, 5};
    else if (tile_bits == 12 && threads == 0) cf = Config{6, 64, 4};
    else
    {
        set_error("layer_plan: tile_bits %d with threads %d is not a tile configuration", tile_bits, threads);
        return E_ARG;
    }
    cf.plan = plan;
    if (nq <= cf.k())
    {
        set_error("layer_plan: %d qubits fit one tile (the small kernel runs them)", nq);
        return E_ARG;
    }
    LayerPlan lp;
    MIPT_HIP_TRY(build_layer_plan(L, nq, closed, grid_x, layer, cf, lp));
    std::vector<int> v;
    const int nb = geo_bond_count(L, grid_x, layer, closed);
    v.push_back(nb);
    for (int b = 0; b < nb; ++b)
    {
        int q0, q1;
        geo_bond(L, grid_x, layer, b, closed, q0, q1);
        v.push_back(q0);
        v.push_back(q1);
    }
    v.push_back(static_cast<int>(lp.passes.size()));
    for (const PassPlan &pp : lp.passes)
    {
        v.insert(v.end(), {pp.k, pp.c, pp.s, pp.nstages});
        for (int si = 0; si < pp.nstages; ++si)
        {
            const StagePlan &sp = pp.st[si];
            v.push_back(sp.use_lds);
            v.push_back(sp.zpair);
            for (int j = 0; j < 2; ++j)
            {
                int q0 = -1, q1 = -1;
                if (sp.bond[j] >= 0) geo_bond(L, grid_x, layer, sp.bond[j], closed, q0, q1);
                v.insert(v.end(), {sp.bond[j], q0, q1});
            }
            v.insert(v.end(), sp.reg_u, sp.reg_u + 6);
            v.insert(v.end(), sp.lane_u, sp.lane_u + 6);
            v.insert(v.end(), sp.wave_u, sp.wave_u + 4);
        }
    }
    *n_out = static_cast<int>(v.size());
    if (static_cast<int>(v.size()) > out_len)
    {
        set_error("layer_plan: %zu ints needed, out_len %d", v.size(), out_len);
        return E_ARG;
    }
    std::copy(v.begin(), v.end(), out);
    return OK;
}

int mipt_hip_batch_set_streams(mipt_hip_batch *b, int streams)
{
    if (!b || (streams != 1 && streams != 2))
    {
        set_error("batch_set_streams: streams must be 1 or 2");
        return E_ARG;
    }
    b->drop_graph();
    b->streams = streams;
    return OK;
}

int mipt_hip_batch_set_profiling(mipt_hip_batch *b, int sample_every)
{
    if (!b || sample_every < 0) { set_error("batch_set_profiling: bad arguments"); return E_ARG; }
    b->prof_every = sample_every;
    b->prof_counter = 0;
    b->ev_used.clear();
    b->gate_launches = 0;
    b->gate_bytes = 0.0;
    return OK;
}

int mipt_hip_batch_pass_stats(mipt_hip_batch *b, double *avg_ms, int *timed, double *bytes_avg, long *launches,
                              double *bytes_total)
{
    if (!b) { set_error("batch_pass_stats: null batch"); return E_ARG; }
    return b->pass_stats(avg_ms, timed, bytes_avg, launches, bytes_total);
}

int mipt_hip_batch_observables(mipt_hip_batch *b, int traj, int tmi_block, double *s_half, double *tmi,
                               double *probe_entropy, double *bloch3)
{
    if (!b || traj < 0 || traj >= static_cast<int>(b->desc.size())) { set_error("batch_observables: bad arguments"); return E_ARG; }
    const TrajDesc &d = b->desc[traj];
    const float2 *psi = b->a.states + d.state_off;
    if (s_half)
    {
        std::vector<int> kept(d.L / 2);
        std::iota(kept.begin(), kept.end(), 0);
        MIPT_HIP_TRY(entropy_bits_f32(b->ew(), psi, d.nq, kept.data(), static_cast<int>(kept.size()), b->stream, s_half));
    }
    if (tmi)
    {
        const int block = tmi_block > 0 ? tmi_block : d.L / 4;
        MIPT_HIP_TRY(tmi_bits_f32(b->ew(), psi, d.nq, d.L, block, 0, b->stream, tmi, nullptr));
    }
    if (probe_entropy || bloch3)
    {
        if (d.probe_site < 0)
        {
            if (probe_entropy) *probe_entropy = NAN;
            if (bloch3) bloch3[0] = bloch3[1] = bloch3[2] = NAN;
        }
        else
        {
            double r[4];
            MIPT_HIP_TRY(rho1_f32(b->ew(), psi, d.nq, d.L, b->stream, r));
            const double tr = r[0] + r[1];
            const double a0 = r[0] / tr, a1 = r[1] / tr, x = r[2] / tr, y = r[3] / tr;
            if (bloch3)
            {
                bloch3[0] = 2 * x;
                bloch3[1] = -2 * y;
                bloch3[2] = a0 - a1;
            }
            if (probe_entropy)
            {
                // cpp/include/mipt/ancilla_mi.hpp:24-52 (nats)
                const double disc = std::min(1.0, std::max(0.0, std::sqrt(std::max(0.0, (a0 - a1) * (a0 - a1) + 4.0 * (x * x + y * y)))));
                const double l0 = 0.5 * (1 + disc), l1 = 0.5 * (1 - disc);
                *probe_entropy = (l0 > 0 ? -l0 * std::log(l0) : 0.0) + (l1 > 0 ? -l1 * std::log(l1) : 0.0);
            }
        }
    }
    return OK;
}

int mipt_hip_batch_observables_all(mipt_hip_batch *b, int tmi_block, int tmi_cycles, double *s_half, double *tmi)
{
    if (!b || tmi_cycles < 0 || (tmi_cycles > 0 && !tmi))
    {
        set_error("batch_observables_all: bad arguments");
        return E_ARG;
    }
    const int count = static_cast<int>(b->desc.size());
    // groups of one (qubits, chain length): every RDM size is shared inside
    std::vector<int> idx(count);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
        return b->desc[x].nq != b->desc[y].nq ? b->desc[x].nq < b->desc[y].nq : b->desc[x].L < b->desc[y].L;
    });
    for (int g0 = 0; g0 < count;)
    {
        int g1 = g0;
        while (g1 < count && b->desc[idx[g1]].nq == b->desc[idx[g0]].nq && b->desc[idx[g1]].L == b->desc[idx[g0]].L) ++g1;
        const int nq = b->desc[idx[g0]].nq, L = b->desc[idx[g0]].L, cnt = g1 - g0;
        std::vector<const float2 *> psis(cnt);
        for (int j = 0; j < cnt; ++j) psis[j] = b->a.states + b->desc[idx[g0 + j]].state_off;
        if (tmi_cycles > 0)
        {
            const int bk = tmi_block > 0 ? tmi_block : L / 4;
            if (L < 4 || bk < 1 || 3 * bk > L || tmi_cycles > L)
            {
                set_error("batch_observables_all: bad TMI block %d for L=%d", bk, L);
                return E_ARG;
            }
            // the seven regions of cpp/include/mipt/tmi/compute.hpp:89-113 per
            // (trajectory, cycle), grouped by size for the batched solves
            struct Item
            {
                int traj_j, cycle, term;
            };
            std::vector<std::pair<int, std::vector<int>>> by_size;  // size -> flattened kept sets
            std::vector<std::vector<Item>> items;
            auto add = [&](int sz, const std::vector<int> &set, Item it) {
                std::size_t k = 0;
                while (k < by_size.size() && by_size[k].first != sz) ++k;
                if (k == by_size.size())
                {
                    by_size.push_back({sz, {}});
                    items.emplace_back();
                }
                by_size[k].second.insert(by_size[k].second.end(), set.begin(), set.end());
                items[k].push_back(it);
            };
            auto range = [&](int begin, int n_) {
                std::vector<int> v;
                for (int i = 0; i < n_; ++i) v.push_back((begin + i) % L);
                return v;
            };
            for (int j = 0; j < cnt; ++j)
                for (int c = 0; c < tmi_cycles; ++c)
                {
                    const auto A = range(c, bk), B = range(c + bk, bk), Cc = range(c + 2 * bk, bk);
                    const auto D = range(c + 3 * bk, L - 3 * bk);
                    auto cat = [](std::vector<int> x, const std::vector<int> &y) {
                        x.insert(x.end(), y.begin(), y.end());
                        return x;
                    };
                    const std::vector<int> sets[7] = {A, B, Cc, D, cat(A, B), cat(A, Cc), cat(B, Cc)};
                    for (int term = 0; term < 7; ++term)
                        add(static_cast<int>(sets[term].size()), sets[term], Item{j, c, term});
                }
            std::vector<double> acc(static_cast<std::size_t>(cnt) * tmi_cycles, 0.0);
            // cycle 0's A u B is the kept list {0, ..., L/2 - 1} of S(L/2) when
            // 2 * block = L / 2: the same RDM, so S(L/2) is taken from it
            const bool reuse_half = s_half && L % 4 == 0 && 2 * bk == L / 2;
            for (std::size_t k = 0; k < by_size.size(); ++k)
            {
                const int sz = by_size[k].first;
                const int ni = static_cast<int>(items[k].size());
                std::vector<const float2 *> ps(ni);
                for (int i = 0; i < ni; ++i) ps[i] = psis[items[k][i].traj_j];
                std::vector<double> sv(ni);
                MIPT_HIP_TRY(entropy_many_bits_f32(b->ew(), ps.data(), nq, by_size[k].second.data(), sz, ni, b->stream, sv.data()));
                for (int i = 0; i < ni; ++i)
                {
                    const Item &it = items[k][i];
                    acc[static_cast<std::size_t>(it.traj_j) * tmi_cycles + it.cycle] += (it.term < 4 ? 1.0 : -1.0) * sv[i];
                    if (reuse_half && it.cycle == 0 && it.term == 4) s_half[idx[g0 + it.traj_j]] = sv[i];
                }
            }
            for (int j = 0; j < cnt; ++j)
                for (int c = 0; c < tmi_cycles; ++c)
                    tmi[static_cast<std::size_t>(idx[g0 + j]) * tmi_cycles + c] = acc[static_cast<std::size_t>(j) * tmi_cycles + c];
        }
        const bool half_done = tmi_cycles > 0 && L % 4 == 0 && 2 * (tmi_block > 0 ? tmi_block : L / 4) == L / 2;
        if (s_half && !half_done)
        {
            std::vector<int> kept(static_cast<std::size_t>(cnt) * (L / 2));
            for (int j = 0; j < cnt; ++j)
                for (int q = 0; q < L / 2; ++q) kept[static_cast<std::size_t>(j) * (L / 2) + q] = q;
            std::vector<double> sv(cnt);
            MIPT_HIP_TRY(entropy_many_bits_f32(b->ew(), psis.data(), nq, kept.data(), L / 2, cnt, b->stream, sv.data()));
            for (int j = 0; j < cnt; ++j) s_half[idx[g0 + j]] = sv[j];
        }
        g0 = g1;
    }
    return OK;
}

int mipt_hip_batch_translation_entropies(mipt_hip_batch *b, int min_kept, int max_kept, int max_s1_kept,
                                         double *s1_nats, double *s2_nats)
{
    if (!b || !s2_nats || min_kept < 1 || max_kept < min_kept || (max_s1_kept >= min_kept && !s1_nats))
    {
        set_error("batch_translation_entropies: bad arguments");
        return E_ARG;
    }
    const int count = static_cast<int>(b->desc.size());
    const int sizes = max_kept - min_kept + 1;
    for (const TrajDesc &d : b->desc)
        if (max_kept >= d.L)
        {
            set_error("batch_translation_entropies: blocks of %d sites in a %d-site chain", max_kept, d.L);
            return E_ARG;
        }
    std::vector<int> idx(count);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
        return b->desc[x].nq != b->desc[y].nq ? b->desc[x].nq < b->desc[y].nq : b->desc[x].L < b->desc[y].L;
    });
    const double ln2 = 0.69314718055994530942;
    for (int g0 = 0; g0 < count;)
    {
        int g1 = g0;
        while (g1 < count && b->desc[idx[g1]].nq == b->desc[idx[g0]].nq && b->desc[idx[g1]].L == b->desc[idx[g0]].L) ++g1;
        const int nq = b->desc[idx[g0]].nq, L = b->desc[idx[g0]].L, cnt = g1 - g0;
        // items (trajectory j, start i) for i = 0..L-1: the cyclic block of
        // cpp/include/mipt/entropy.hpp:521-577, averaged over the L starts
        std::vector<const float2 *> ps(static_cast<std::size_t>(cnt) * L);
        for (int j = 0; j < cnt; ++j)
            for (int i = 0; i < L; ++i) ps[static_cast<std::size_t>(j) * L + i] = b->a.states + b->desc[idx[g0 + j]].state_off;
        for (int l = min_kept; l <= max_kept; ++l)
        {
            std::vector<int> kept(static_cast<std::size_t>(cnt) * L * l);
            for (int j = 0; j < cnt; ++j)
                for (int i = 0; i < L; ++i)
                    for (int k = 0; k < l; ++k) kept[(static_cast<std::size_t>(j) * L + i) * l + k] = (i + k) % L;
            std::vector<double> v(static_cast<std::size_t>(cnt) * L);
            MIPT_HIP_TRY(purity_many_f32(b->ew(), ps.data(), nq, kept.data(), l, cnt * L, b->stream, v.data()));
            for (int j = 0; j < cnt; ++j)
            {
                double acc = 0.0;
                for (int i = 0; i < L; ++i)
                {
                    const double pu = v[static_cast<std::size_t>(j) * L + i];
                    acc += (pu > 0.0 && pu < 1.0) ? -std::log(pu) : 0.0;
                }
                s2_nats[static_cast<std::size_t>(idx[g0 + j]) * sizes + (l - min_kept)] = acc / L;
            }
            if (s1_nats)
            {
                if (l <= max_s1_kept)
                {
                    MIPT_HIP_TRY(entropy_many_bits_f32(b->ew(), ps.data(), nq, kept.data(), l, cnt * L, b->stream, v.data()));
                    for (int j = 0; j < cnt; ++j)
                    {
                        double acc = 0.0;
                        for (int i = 0; i < L; ++i) acc += v[static_cast<std::size_t>(j) * L + i];
                        s1_nats[static_cast<std::size_t>(idx[g0 + j]) * sizes + (l - min_kept)] = acc * ln2 / L;
                    }
                }
                else
                    for (int j = 0; j < cnt; ++j)
                        s1_nats[static_cast<std::size_t>(idx[g0 + j]) * sizes + (l - min_kept)] = NAN;
            }
        }
        g0 = g1;
    }
    return OK;
}

int mipt_hip_batch_tmi(mipt_hip_batch *b, int traj, int block, int cycle, int order, double *tmi)
{
    if (!b || traj < 0 || traj >= static_cast<int>(b->desc.size()) || !tmi || (order != 1 && order != 2))
    {
        set_error("batch_tmi: bad arguments");
        return E_ARG;
    }
    const TrajDesc &d = b->desc[traj];
    const int blk = block > 0 ? block : d.L / 4;
    return tmi_bits_f32(b->ew(), b->a.states + d.state_off, d.nq, d.L, blk, cycle, b->stream, tmi, nullptr, order);
}

} // extern "C"
