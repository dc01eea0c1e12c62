// CPU FP64 (complex128) restatement of the amarce4/mipt qubit MIPT trajectory
// path.  TEST INFRASTRUCTURE ONLY: this file is the checker for the HIP engine
// in mipt_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and never as the thing measured or shipped.
//
// Every function cites the reference file:line whose behaviour it restates
// (paths relative to the reference repository root).  Functions whose exact
// behaviour nothing in the reference or mathematics pins are marked
// "parity unpinned" here; everything else is pinned by tests/test_oracle.py.
//
// State layout: complex128 interleaved (re, im), qubit q <-> index bit q,
// little-endian (cpp/include/mipt/cusv/matrix.hpp:5-9).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

namespace
{
using cd = std::complex<double>;

// ---------------------------------------------------------------- seeding
// cpp/include/mipt/types.hpp:262-273
std::uint64_t splitmix64(std::uint64_t v)
{
    v += 0x9e3779b97f4a7c15ull;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
    return v ^ (v >> 31);
}

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  The counter-based
// stream the batched engine draws from; pinned by the published Random123
// known-answer vectors in tests/test_oracle.py.
void philox4x32_10(const std::uint32_t key_in[2], const std::uint32_t ctr_in[4], std::uint32_t out[4])
{
    std::uint32_t k0 = key_in[0], k1 = key_in[1];
    std::uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    for (int round = 0; round < 10; ++round)
    {
        const std::uint64_t p0 = std::uint64_t{0xD2511F53u} * c0;
        const std::uint64_t p1 = std::uint64_t{0xCD9E8D57u} * c2;
        const std::uint32_t hi0 = static_cast<std::uint32_t>(p0 >> 32), lo0 = static_cast<std::uint32_t>(p0);
        const std::uint32_t hi1 = static_cast<std::uint32_t>(p1 >> 32), lo1 = static_cast<std::uint32_t>(p1);
        const std::uint32_t n0 = hi1 ^ c1 ^ k0;
        const std::uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Stream identifiers (counter word 3) of the engine's counter-based draws.
constexpr std::uint32_t STREAM_GATE = 0x47415445u;  // gate normals
constexpr std::uint32_t STREAM_FLAG = 0x464c4147u;  // Bernoulli(p) flags
constexpr std::uint32_t STREAM_MEAS = 0x4d454153u;  // measurement uniforms

// 53-bit uniform in [0,1) from two 32-bit words.
double u53(std::uint32_t a, std::uint32_t b)
{
    return (static_cast<double>(a >> 5) * 67108864.0 + static_cast<double>(b >> 6)) * (1.0 / 9007199254740992.0);
}

void philox_key(std::uint64_t seed, std::uint32_t key[2])
{
    key[0] = static_cast<std::uint32_t>(seed & 0xffffffffu);
    key[1] = static_cast<std::uint32_t>(seed >> 32);
}

// Two uniforms (u, v) for draw `slot` of (layer, item, stream).
void philox_uniform_pair(std::uint64_t seed, std::uint32_t layer, std::uint32_t item,
                         std::uint32_t slot, std::uint32_t stream, double &u, double &v)
{
    std::uint32_t key[2];
    philox_key(seed, key);
    const std::uint32_t ctr[4] = {slot, item, layer, stream};
    std::uint32_t r[4];
    philox4x32_10(key, ctr, r);
    u = u53(r[0], r[1]);
    v = u53(r[2], r[3]);
}

// ---------------------------------------------------------------- Haar U(4)
// cpp/include/mipt/circuits/haar.hpp:39-85: four columns of complex standard
// normals, modified Gram-Schmidt in column order, a column redrawn when its
// residual norm is below 1e-14.  Row-major, basis index 2*b(q1) + b(q0).
template <typename NormalSource>
void haar_u4_gram_schmidt(NormalSource &&normal, cd q[16])
{
    for (int col = 0; col < 4; ++col)
    {
        cd v[4];
        for (int row = 0; row < 4; ++row)
        {
            // Re is drawn before Im (the left-to-right order of a clang build;
            // the reference leaves the order to the compiler, haar.hpp:51).
            const double re = normal();
            const double im = normal();
            v[row] = cd(re, im);
        }
        for (int prev = 0; prev < col; ++prev)
        {
            cd dot = 0.0;
            for (int row = 0; row < 4; ++row) dot += std::conj(q[row * 4 + prev]) * v[row];
            for (int row = 0; row < 4; ++row) v[row] -= q[row * 4 + prev] * dot;
        }
        double norm2 = 0.0;
        for (int row = 0; row < 4; ++row) norm2 += std::norm(v[row]);
        const double norm = std::sqrt(norm2);
        if (norm < 1e-14)
        {
            --col;
            continue;
        }
        for (int row = 0; row < 4; ++row) q[row * 4 + col] = v[row] / norm;
    }
}

// ---------------------------------------------------------------- layers
// cpp/include/mipt/circuits/haar.hpp:226-243: layer t starts at t mod 2; with
// `closed` an odd layer ends with the wrap bond (n-1, 0), q0 = n-1.
int layer_bonds(int n, int start, int closed, int *q0, int *q1)
{
    const int stop = (closed && start == 1 && n > 2) ? n : (n - 1);
    int b = 0;
    for (int i = start; i < stop; i += 2)
    {
        if (q0) q0[b] = i;
        if (q1) q1[b] = (i + 1) % n;
        ++b;
    }
    return b;
}

// ---------------------------------------------------------------- gate apply
// Applies a dense 2^k x 2^k row-major matrix on `targets` (bit j of the local
// index = targets[j]), the semantics of a row-layout custatevecApplyMatrix
// call (cpp/include/mipt/cusv/engine.hpp:236-277).
void apply_dense(cd *psi, int n, const int *targets, int k, const cd *m)
{
    const std::uint64_t dim = std::uint64_t{1} << n;
    const int d = 1 << k;
    std::uint64_t tmask = 0;
    std::vector<std::uint64_t> off(d, 0);
    for (int j = 0; j < k; ++j) tmask |= std::uint64_t{1} << targets[j];
    for (int a = 0; a < d; ++a)
        for (int j = 0; j < k; ++j)
            if ((a >> j) & 1) off[a] |= std::uint64_t{1} << targets[j];
    std::vector<cd> in(d), out(d);
    for (std::uint64_t base = 0; base < dim; ++base)
    {
        if (base & tmask) continue;
        for (int a = 0; a < d; ++a) in[a] = psi[base | off[a]];
        for (int r = 0; r < d; ++r)
        {
            cd acc = 0.0;
            for (int c = 0; c < d; ++c) acc += m[r * d + c] * in[c];
            out[r] = acc;
        }
        for (int a = 0; a < d; ++a) psi[base | off[a]] = out[a];
    }
}

// Two-qubit specialisation of apply_dense (same arithmetic order).
void apply_bond(cd *psi, int n, int q0, int q1, const cd *u)
{
    const std::uint64_t m0 = std::uint64_t{1} << q0, m1 = std::uint64_t{1} << q1;
    const int lo = std::min(q0, q1), hi = std::max(q0, q1);
    const std::uint64_t groups = std::uint64_t{1} << (n - 2);
    for (std::uint64_t j = 0; j < groups; ++j)
    {
        // j with zero bits inserted at positions lo and hi: the group bases in
        // increasing order (each amplitude group visited once, as before)
        std::uint64_t base = ((j >> lo) << (lo + 1)) | (j & ((std::uint64_t{1} << lo) - 1));
        base = ((base >> hi) << (hi + 1)) | (base & ((std::uint64_t{1} << hi) - 1));
        const cd in[4] = {psi[base], psi[base | m0], psi[base | m1], psi[base | m0 | m1]};
        cd out[4];
        for (int r = 0; r < 4; ++r)
        {
            cd acc = 0.0;
            for (int c = 0; c < 4; ++c) acc += u[r * 4 + c] * in[c];
            out[r] = acc;
        }
        psi[base] = out[0];
        psi[base | m0] = out[1];
        psi[base | m1] = out[2];
        psi[base | m0 | m1] = out[3];
    }
}

// ---------------------------------------------------------------- measurement
// cpp/include/mipt/free_energy_measure.hpp:19-95: one site, long-double norms,
// outcome = 1 iff u < P(1), collapse scaled by 1/sqrt(raw branch norm).
// Returns 0 or a negative status matching the device C ABI's -1..-4
// (cpp/src/cuda/free_energy_measure.cu:137-238).
int measure_collapse(cd *psi, int n, int q, double u, int *outcome, double *prob)
{
    if (psi == nullptr || n < 1 || n >= 63 || q < 0 || q >= n || !(u >= 0.0 && u < 1.0)) return -1;
    const std::uint64_t dim = std::uint64_t{1} << n;
    const std::uint64_t mask = std::uint64_t{1} << q;
    long double total = 0.0L, one = 0.0L;
    for (std::uint64_t i = 0; i < dim; ++i)
    {
        const long double w = static_cast<long double>(std::norm(psi[i]));
        total += w;
        if (i & mask) one += w;
    }
    const double t = static_cast<double>(total);
    const double b1 = static_cast<double>(one);
    if (!(t > 0.0) || !std::isfinite(t) || !std::isfinite(b1) || b1 < 0.0 || b1 > t * (1.0 + 1.0e-10)) return -2;
    const double p1 = std::clamp(b1 / t, 0.0, 1.0);
    const int o = u < p1 ? 1 : 0;
    const double cp = o ? p1 : 1.0 - p1;
    const double branch = o ? b1 : t - b1;
    if (!(cp > 0.0) || !std::isfinite(cp)) return -3;
    if (!(branch > 0.0) || !std::isfinite(branch)) return -4;
    const double inv = 1.0 / std::sqrt(branch);
    for (std::uint64_t i = 0; i < dim; ++i)
    {
        const int bit = (i & mask) ? 1 : 0;
        psi[i] = (bit == o) ? psi[i] * inv : cd(0.0, 0.0);
    }
    *outcome = o;
    *prob = cp;
    return 0;
}

// Joint chunked sampling of one layer's sites, the cuStateVec engine's
// semantics (cpp/include/mipt/cusv/engine.hpp:293-368): per chunk of at most
// `chunk_bits` ascending sites, an FP64 histogram whose bin bit i is
// sites[i]; draw = u * total; the chosen bin is the first whose running sum
// exceeds draw (a starved bin falls back to the heaviest); collapse by
// 1/sqrt(bin weight).  One uniform per chunk.  Returns the surprisal.
double measure_layer_joint(cd *psi, int n, const int *sites, int count, int chunk_bits,
                           const double *uniforms, int *outcomes, int *status)
{
    *status = 0;
    const std::uint64_t dim = std::uint64_t{1} << n;
    double surprisal = 0.0;
    int chunk = 0;
    for (int begin = 0; begin < count; begin += chunk_bits, ++chunk)
    {
        const int end = std::min(count, begin + chunk_bits);
        const int m = end - begin;
        const std::size_t bins = std::size_t{1} << m;
        std::vector<double> h(bins, 0.0);
        for (std::uint64_t i = 0; i < dim; ++i)
        {
            std::size_t b = 0;
            for (int j = 0; j < m; ++j) b |= static_cast<std::size_t>((i >> sites[begin + j]) & 1u) << j;
            h[b] += std::norm(psi[i]);
        }
        double total = 0.0;
        for (double v : h) total += v;
        if (!(total > 0.0) || !std::isfinite(total)) { *status = -2; return surprisal; }
        const double draw = uniforms[chunk] * total;
        double running = 0.0;
        std::size_t chosen = bins - 1;
        for (std::size_t i = 0; i < bins; ++i)
        {
            running += h[i];
            if (draw < running) { chosen = i; break; }
        }
        double branch = h[chosen];
        if (!(branch > 0.0))
        {
            chosen = static_cast<std::size_t>(std::max_element(h.begin(), h.end()) - h.begin());
            branch = h[chosen];
            if (!(branch > 0.0)) { *status = -3; return surprisal; }
        }
        std::uint64_t cmask = 0, cval = 0;
        for (int j = 0; j < m; ++j)
        {
            const int bit = static_cast<int>((chosen >> j) & 1u);
            outcomes[begin + j] = bit;
            cmask |= std::uint64_t{1} << sites[begin + j];
            if (bit) cval |= std::uint64_t{1} << sites[begin + j];
        }
        surprisal += -std::log(branch / total);
        const double inv = 1.0 / std::sqrt(branch);
        for (std::uint64_t i = 0; i < dim; ++i) psi[i] = ((i & cmask) == cval) ? psi[i] * inv : cd(0.0, 0.0);
    }
    return surprisal;
}

// ---------------------------------------------------------------- entropy
// cpp/include/mipt/tmi/linalg.hpp:125-141 (cut-offs).
constexpr double LOG2_EPS = 1.0e-15;

double entropy_term_bits(double lambda) { return lambda <= LOG2_EPS ? 0.0 : -lambda * std::log2(lambda); }
double finish_entropy(double s) { return std::abs(s) < 1.0e-12 ? 0.0 : s; }

// Cyclic complex Jacobi eigenvalues of a Hermitian matrix (row-major),
// the algorithm of cpp/include/mipt/util/spectral.hpp:40-160 (hermitize, then
// sweeps of 2x2 rotations until the off-diagonal mass is below tolerance).
int jacobi_eigvalsh(std::vector<cd> a, int n, double *w)
{
    auto at = [&](int r, int c) -> cd & { return a[static_cast<std::size_t>(r) * n + c]; };
    double scale = 0.0;
    for (int r = 0; r < n; ++r)
    {
        at(r, r) = cd(at(r, r).real(), 0.0);
        scale = std::max(scale, std::abs(at(r, r).real()));
        for (int c = r + 1; c < n; ++c)
        {
            const cd mean = 0.5 * (at(r, c) + std::conj(at(c, r)));
            at(r, c) = mean;
            at(c, r) = std::conj(mean);
            scale = std::max(scale, std::abs(mean));
        }
    }
    if (n == 1) { w[0] = at(0, 0).real(); return 0; }
    const double tol = 1.0e-15 * std::max(scale, 1e-300);
    int converged = 0;
    for (int sweep = 0; sweep < 60; ++sweep)
    {
        double off = 0.0;
        for (int r = 0; r < n; ++r)
            for (int c = r + 1; c < n; ++c) off = std::max(off, std::abs(at(r, c)));
        if (off <= tol) { converged = 1; break; }
        for (int p = 0; p < n - 1; ++p)
        {
            for (int q = p + 1; q < n; ++q)
            {
                const cd apq = at(p, q);
                const double b = std::abs(apq);
                if (b <= 0.0) continue;
                const double app = at(p, p).real(), aqq = at(q, q).real();
                const double tau = (aqq - app) / (2.0 * b);
                const double t = std::copysign(1.0 / (std::abs(tau) + std::sqrt(1.0 + tau * tau)), tau);
                const double c = 1.0 / std::sqrt(1.0 + t * t);
                const double s = t * c;
                const cd ph = apq / b;
                for (int k = 0; k < n; ++k)
                {
                    if (k == p || k == q) continue;
                    const cd kp = at(k, p), kq = at(k, q);
                    const cd nkp = c * kp * ph - s * kq;
                    const cd nkq = s * kp * ph + c * kq;
                    at(k, p) = nkp; at(p, k) = std::conj(nkp);
                    at(k, q) = nkq; at(q, k) = std::conj(nkq);
                }
                at(p, p) = cd(app - t * b, 0.0);
                at(q, q) = cd(aqq + t * b, 0.0);
                at(p, q) = 0.0;
                at(q, p) = 0.0;
            }
        }
    }
    for (int i = 0; i < n; ++i) w[i] = at(i, i).real();
    return converged ? 0 : 1;
}

// Reduced density matrix of `kept` (bit j of the row index = kept[j]; the
// environment is the complement in ascending order), FP64, row-major:
// cpp/include/mipt/tmi/linalg.hpp:885-903 (basis_offsets) and 1097-1142
// (reduced_density_manual_from_state).
void rdm(const cd *psi, int n, const int *kept, int count, cd *rho)
{
    std::vector<int> env;
    std::uint64_t km = 0;
    for (int j = 0; j < count; ++j) km |= std::uint64_t{1} << kept[j];
    for (int q = 0; q < n; ++q) if (!((km >> q) & 1)) env.push_back(q);
    const std::size_t rows = std::size_t{1} << count;
    const std::size_t cols = std::size_t{1} << env.size();
    std::vector<std::uint64_t> roff(rows, 0), eoff(cols, 0);
    for (std::size_t r = 0; r < rows; ++r)
        for (int j = 0; j < count; ++j) if ((r >> j) & 1) roff[r] |= std::uint64_t{1} << kept[j];
    for (std::size_t c = 0; c < cols; ++c)
        for (std::size_t j = 0; j < env.size(); ++j) if ((c >> j) & 1) eoff[c] |= std::uint64_t{1} << env[j];
    std::fill(rho, rho + rows * rows, cd(0.0, 0.0));
    std::vector<cd> col(rows);
    for (std::size_t c = 0; c < cols; ++c)
    {
        for (std::size_t r = 0; r < rows; ++r) col[r] = psi[eoff[c] | roff[r]];
        for (std::size_t r = 0; r < rows; ++r)
        {
            const cd a = col[r];
            if (a == cd(0.0, 0.0)) continue;
            for (std::size_t s = 0; s < rows; ++s) rho[r * rows + s] += a * std::conj(col[s]);
        }
    }
}
} // namespace

// =================================================================== C ABI
// Haar U(4) of (layer, bond) on the counter stream: Box-Muller normals (cos
// then sin of each uniform pair, Re before Im), Gram-Schmidt as above.
void philox_haar(std::uint64_t seed, std::uint32_t t, std::uint32_t b, cd q[16])
{
    std::uint32_t slot = 0;
    int have = 0;
    double saved = 0.0;
    auto normal = [&]() {
        if (have) { have = 0; return saved; }
        double u, v;
        philox_uniform_pair(seed, t, b, slot++, STREAM_GATE, u, v);
        const double r = std::sqrt(-2.0 * std::log(1.0 - u));
        const double a = 6.283185307179586476925286766559 * v;
        saved = r * std::sin(a);
        have = 1;
        return r * std::cos(a);
    };
    haar_u4_gram_schmidt(normal, q);
}

extern "C" {

// One gate of the counter stream, complex128 row-major interleaved (32 doubles).
void orc_philox_haar(std::uint64_t seed, std::uint32_t layer, std::uint32_t bond, double *out)
{
    cd q[16];
    philox_haar(seed, layer, bond, q);
    for (int e = 0; e < 16; ++e)
    {
        out[2 * e] = q[e].real();
        out[2 * e + 1] = q[e].imag();
    }
}

std::uint64_t orc_splitmix64(std::uint64_t v) { return splitmix64(v); }

// cpp/include/mipt/types.hpp:270-273
std::uint64_t orc_trajectory_seed(std::uint64_t master, std::uint64_t index)
{
    return splitmix64(master + 0x2545f4914f6cdd1dull * (index + 1ull));
}

// cpp/include/mipt/types.hpp:23-38
void orc_linspace(double first, double last, int count, double *out)
{
    if (count <= 0) return;
    if (count == 1) { out[0] = first; return; }
    const double step = (last - first) / static_cast<double>(count - 1);
    for (int i = 0; i < count; ++i) out[i] = first + step * static_cast<double>(i);
    out[count - 1] = last;
}

void orc_philox4x32_10(const std::uint32_t *key, const std::uint32_t *ctr, std::uint32_t *out)
{
    philox4x32_10(key, ctr, out);
}

// First uniform of draw `slot` of (layer, item, stream) on the counter stream.
double orc_philox_uniform(std::uint64_t seed, std::uint32_t layer, std::uint32_t item, std::uint32_t slot,
                          std::uint32_t stream)
{
    double u, v;
    philox_uniform_pair(seed, layer, item, slot, stream, u, v);
    return u;
}

int orc_layer_bonds(int n, int layer, int closed, int *q0, int *q1)
{
    return layer_bonds(n, layer & 1, closed, q0, q1);
}

// Total bonds over `layers` layers (gate-buffer sizing).
int orc_total_bonds(int n, int layers, int closed)
{
    int total = 0;
    for (int t = 0; t < layers; ++t) total += layer_bonds(n, t & 1, closed, nullptr, nullptr);
    return total;
}

// Reference-compatible host stream (parity unpinned at the bit level: the
// reference's mt19937 -> normal/bernoulli streams are implementation-defined,
// and its nvq++ build uses a different standard library).  Follows
// cpp/include/mipt/circuit.hpp:129-140 (seeding) and
// cpp/include/mipt/circuits/haar.hpp:226-305 (draw order: a layer's gates in
// bond order, then n Bernoulli(p) flags; none drawn for p<=0 or p>=1).
// gates: [sum of bonds][16 complex] interleaved; flags: [layers][n];
// uniforms: [layers][n] one measurement uniform per measured site, drawn
// ascending from the measurement stream (0 for unmeasured sites).
void orc_mt_build_layers(int n, int layers, double p, int closed, std::uint64_t seed,
                         double *gates, std::uint8_t *flags, double *uniforms)
{
    std::seed_seq seq{static_cast<std::uint32_t>(seed & 0xffffffffu), static_cast<std::uint32_t>(seed >> 32)};
    std::mt19937 rng(seq);
    const std::uint64_t ms = splitmix64(seed ^ 0xa0761d0dull);
    std::seed_seq mseq{static_cast<std::uint32_t>(ms & 0xffffffffu), static_cast<std::uint32_t>(ms >> 32)};
    std::mt19937 mrng(mseq);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    std::size_t g = 0;
    for (int t = 0; t < layers; ++t)
    {
        const int nb = layer_bonds(n, t & 1, closed, nullptr, nullptr);
        for (int b = 0; b < nb; ++b, ++g)
        {
            std::normal_distribution<double> normal(0.0, 1.0);
            cd q[16];
            haar_u4_gram_schmidt([&]() { return normal(rng); }, q);
            for (int e = 0; e < 16; ++e)
            {
                gates[g * 32 + 2 * e] = q[e].real();
                gates[g * 32 + 2 * e + 1] = q[e].imag();
            }
        }
        std::uint8_t *f = flags + static_cast<std::size_t>(t) * n;
        if (p <= 0.0) std::fill(f, f + n, 0);
        else if (p >= 1.0) std::fill(f, f + n, 1);
        else
        {
            std::bernoulli_distribution md(p);
            for (int s = 0; s < n; ++s) f[s] = md(rng) ? 1 : 0;
        }
        double *u = uniforms + static_cast<std::size_t>(t) * n;
        for (int s = 0; s < n; ++s) u[s] = f[s] ? uni(mrng) : 0.0;
    }
}

// Counter-based stream of the batched engine (the same draws the device
// generator makes): gate normals by Box-Muller on Philox uniforms keyed by
// (seed, layer, bond); flags u < p keyed by (seed, layer, site); measurement
// uniforms keyed by (seed, layer, site).  Same outputs as orc_mt_build_layers.
void orc_philox_build_layers(int n, int layers, double p, int closed, std::uint64_t seed,
                             double *gates, std::uint8_t *flags, double *uniforms)
{
    std::size_t g = 0;
    for (int t = 0; t < layers; ++t)
    {
        const int nb = layer_bonds(n, t & 1, closed, nullptr, nullptr);
        for (int b = 0; b < nb; ++b, ++g)
        {
            cd q[16];
            philox_haar(seed, static_cast<std::uint32_t>(t), static_cast<std::uint32_t>(b), q);
            for (int e = 0; e < 16; ++e)
            {
                gates[g * 32 + 2 * e] = q[e].real();
                gates[g * 32 + 2 * e + 1] = q[e].imag();
            }
        }
        std::uint8_t *f = flags + static_cast<std::size_t>(t) * n;
        double *u = uniforms + static_cast<std::size_t>(t) * n;
        for (int s = 0; s < n; ++s)
        {
            double a, b2, m, unused;
            philox_uniform_pair(seed, static_cast<std::uint32_t>(t), static_cast<std::uint32_t>(s), 0, STREAM_FLAG, a, b2);
            f[s] = (p <= 0.0) ? 0 : (p >= 1.0) ? 1 : (a < p ? 1 : 0);
            philox_uniform_pair(seed, static_cast<std::uint32_t>(t), static_cast<std::uint32_t>(s), 0, STREAM_MEAS, m, unused);
            u[s] = f[s] ? m : 0.0;
        }
    }
}

void orc_apply_dense(double *psi, int n, const int *targets, int k, const double *m)
{
    apply_dense(reinterpret_cast<cd *>(psi), n, targets, k, reinterpret_cast<const cd *>(m));
}

int orc_measure_collapse(double *psi, int n, int q, double u, int *outcome, double *prob)
{
    return measure_collapse(reinterpret_cast<cd *>(psi), n, q, u, outcome, prob);
}

double orc_measure_layer_joint(double *psi, int n, const int *sites, int count, int chunk_bits,
                               const double *uniforms, int *outcomes, int *status)
{
    return measure_layer_joint(reinterpret_cast<cd *>(psi), n, sites, count, chunk_bits, uniforms, outcomes, status);
}

// One trajectory's layers from the given state (usually |0...0>), the loop of
// cpp/include/mipt/circuit.hpp:240-282: the layer's bond gates in bond order,
// then its measured sites.  mode 0: sequential per-site sampling with the
// per-site uniforms (free_energy_measure.hpp semantics; p1[t*n+s] receives
// the conditional P(1) each measured site saw, outcomes[t*n+s] its outcome,
// -1 for unmeasured).  mode 1: joint chunked sampling (engine.hpp), using
// uniforms[t*n + chunk] as the chunk uniforms.  Returns 0 or a status.
int orc_run_layers(double *psi_ri, int n, int layer_begin, int layer_end, int closed, const double *gates,
                   const std::uint8_t *flags, const double *uniforms, int mode, int chunk_bits,
                   int *outcomes, double *p1)
{
    cd *psi = reinterpret_cast<cd *>(psi_ri);
    const cd *g = reinterpret_cast<const cd *>(gates);
    std::size_t gi = 0;
    for (int t = 0; t < layer_begin; ++t) gi += layer_bonds(n, t & 1, closed, nullptr, nullptr);
    std::vector<int> q0(n), q1(n), sites;
    for (int t = layer_begin; t < layer_end; ++t)
    {
        const int nb = layer_bonds(n, t & 1, closed, q0.data(), q1.data());
        for (int b = 0; b < nb; ++b, ++gi) apply_bond(psi, n, q0[b], q1[b], g + gi * 16);
        sites.clear();
        for (int s = 0; s < n; ++s)
        {
            outcomes[static_cast<std::size_t>(t) * n + s] = -1;
            p1[static_cast<std::size_t>(t) * n + s] = -1.0;
            if (flags[static_cast<std::size_t>(t) * n + s]) sites.push_back(s);
        }
        if (sites.empty()) continue;
        if (mode == 0)
        {
            for (int s : sites)
            {
                // P(1) before the collapse, for near-threshold flagging.
                const std::uint64_t dim = std::uint64_t{1} << n;
                long double tot = 0.0L, one = 0.0L;
                for (std::uint64_t i = 0; i < dim; ++i)
                {
                    const long double w = std::norm(psi[i]);
                    tot += w;
                    if ((i >> s) & 1) one += w;
                }
                int o;
                double prob;
                const int st = measure_collapse(psi, n, s, uniforms[static_cast<std::size_t>(t) * n + s], &o, &prob);
                if (st != 0) return st;
                outcomes[static_cast<std::size_t>(t) * n + s] = o;
                p1[static_cast<std::size_t>(t) * n + s] = static_cast<double>(one / tot);
            }
        }
        else
        {
            std::vector<int> o(sites.size());
            int st = 0;
            measure_layer_joint(psi, n, sites.data(), static_cast<int>(sites.size()), chunk_bits,
                                uniforms + static_cast<std::size_t>(t) * n, o.data(), &st);
            if (st != 0) return st;
            for (std::size_t j = 0; j < sites.size(); ++j) outcomes[static_cast<std::size_t>(t) * n + sites[j]] = o[j];
        }
    }
    return 0;
}

void orc_rdm(const double *psi, int n, const int *kept, int count, double *rho_ri)
{
    rdm(reinterpret_cast<const cd *>(psi), n, kept, count, reinterpret_cast<cd *>(rho_ri));
}

int orc_jacobi_eigvalsh(const double *a_ri, int n, double *w)
{
    const cd *a = reinterpret_cast<const cd *>(a_ri);
    return jacobi_eigvalsh(std::vector<cd>(a, a + static_cast<std::size_t>(n) * n), n, w);
}

// -sum lambda log2 lambda with the reference cut-offs (linalg.hpp:125-141).
double orc_entropy_bits_from_eigs(const double *w, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += entropy_term_bits(w[i]);
    return finish_entropy(s);
}

// Probe qubit 2x2 RDM with FP64 accumulation: out = {rho00, rho11, Re rho01,
// Im rho01}, rho01 = sum psi[..0..] conj(psi[..1..])
// (cpp/src/cuda/ancilla_rdm.cu:228-297, qubit trace).
void orc_rho1(const double *psi_ri, int nq, int qubit, double *out)
{
    const cd *psi = reinterpret_cast<const cd *>(psi_ri);
    const std::uint64_t env = std::uint64_t{1} << (nq - 1);
    const std::uint64_t low = (std::uint64_t{1} << qubit) - 1;
    double r00 = 0, r11 = 0, re = 0, im = 0;
    for (std::uint64_t e = 0; e < env; ++e)
    {
        const std::uint64_t i0 = (e & low) | ((e & ~low) << 1);
        const std::uint64_t i1 = i0 | (std::uint64_t{1} << qubit);
        const cd a = psi[i0], b = psi[i1];
        r00 += std::norm(a);
        r11 += std::norm(b);
        const cd x = a * std::conj(b);
        re += x.real();
        im += x.imag();
    }
    out[0] = r00; out[1] = r11; out[2] = re; out[3] = im;
}

// Entropy in nats of a 2x2 probe RDM (cpp/include/mipt/ancilla_mi.hpp:24-52).
double orc_rho1_entropy_nats(const double *r)
{
    const double tr = r[0] + r[1];
    if (!(tr > 0.0) || !std::isfinite(tr)) return NAN;
    const double a = r[0] / tr, d = r[1] / tr, x = r[2] / tr, y = r[3] / tr;
    const double disc = std::clamp(std::sqrt(std::max(0.0, (a - d) * (a - d) + 4.0 * (x * x + y * y))), 0.0, 1.0);
    const double l0 = 0.5 * (1.0 + disc), l1 = 0.5 * (1.0 - disc);
    auto term = [](double l) { return l > 0.0 ? -l * std::log(l) : 0.0; };
    return term(l0) + term(l1);
}

} // extern "C"
