// Information measures of the references' joint density matrix (host side,
// FP64): the one-, two-, three- and four-reference samples of
// cpp/include/mipt/probed/kernels.hpp:106-147, in nats.
//
//   k = 1  S(A)                                      ancilla_mi.hpp:24-52
//   k = 2  S(AB), I(A:B), negativity, log-negativity ancilla_mi.hpp:222-262
//   k = 3  tripartite information, mean pair MI, min cut negativity, joint
//          and mean single purity                    ancilla_info.hpp:46-66,
//                                                    analysis/cut_negativity.hpp:139-200
//   k = 4  S(ABCD), mean I2 over pairs, mean I3 over triples, I4
//                                                    ancilla_info.hpp:75-139
//
// Eigenvalues come from cyclic Jacobi rotations on the real symmetric
// embedding [[Re, -Im], [Im, Re]] of the Hermitian part, which holds every
// eigenvalue twice (Golub & Van Loan, Matrix Computations, 8.5).
#include "mipt_hip.h"

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace
{
using cplx = std::complex<double>;

// eigenvalues (each once, ascending) of the Hermitian part of a D x D matrix
bool hermitian_eigenvalues(const std::vector<cplx> &m, int D, std::vector<double> &w)
{
    const int n = 2 * D;
    std::vector<double> a(static_cast<std::size_t>(n) * n);
    auto at = [&](int r, int c) -> double & { return a[static_cast<std::size_t>(r) * n + c]; };
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c)
        {
            const cplx h = 0.5 * (m[r * D + c] + std::conj(m[c * D + r]));
            at(r, c) = at(r + D, c + D) = h.real();
            at(r, c + D) = -h.imag();
            at(r + D, c) = h.imag();
        }
    bool done = false;
    for (int sweep = 0; sweep < 100 && !done; ++sweep)
    {
        double off = 0.0, scale = 0.0;
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) (r == c ? scale : off) += at(r, c) * at(r, c);
        if (off <= 1e-30 * std::max(scale, 1e-300))
        {
            done = true;
            break;
        }
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q)
            {
                const double apq = at(p, q);
                if (apq == 0.0) continue;
                // rotation angle zeroing (p, q): tan(2 theta) = 2 apq / (aqq - app)
                const double theta = 0.5 * (at(q, q) - at(p, p)) / apq;
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k)
                {
                    const double kp = at(k, p), kq = at(k, q);
                    at(k, p) = c * kp - s * kq;
                    at(k, q) = s * kp + c * kq;
                }
                for (int k = 0; k < n; ++k)
                {
                    const double pk = at(p, k), qk = at(q, k);
                    at(p, k) = c * pk - s * qk;
                    at(q, k) = s * pk + c * qk;
                }
                at(p, q) = at(q, p) = 0.0;
            }
    }
    if (!done) return false;
    std::vector<double> all(n);
    for (int i = 0; i < n; ++i) all[i] = at(i, i);
    std::sort(all.begin(), all.end());
    w.resize(D);
    for (int i = 0; i < D; ++i) w[i] = 0.5 * (all[2 * i] + all[2 * i + 1]);
    return true;
}

// reduced matrix of the references in `mask` (bit j = reference j), local
// bit order kept, normalised to trace 1
std::vector<cplx> partial_trace(const std::vector<cplx> &rho, int k, unsigned mask, int &d)
{
    std::vector<int> keep, drop;
    for (int j = 0; j < k; ++j) (((mask >> j) & 1u) ? keep : drop).push_back(j);
    d = 1 << keep.size();
    const int de = 1 << drop.size(), D = 1 << k;
    auto compose = [&](int sub, int env) {
        int x = 0;
        for (std::size_t i = 0; i < keep.size(); ++i) x |= ((sub >> i) & 1) << keep[i];
        for (std::size_t i = 0; i < drop.size(); ++i) x |= ((env >> i) & 1) << drop[i];
        return x;
    };
    std::vector<cplx> out(static_cast<std::size_t>(d) * d);
    double tr = 0.0;
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c)
        {
            cplx v = 0.0;
            for (int e = 0; e < de; ++e) v += rho[compose(r, e) * D + compose(c, e)];
            out[r * d + c] = v;
        }
    for (int r = 0; r < d; ++r) tr += out[r * d + r].real();
    for (auto &v : out) v /= tr;
    return out;
}

// von Neumann entropy (nats); eigenvalues below -2e-5 are an error, the rest
// are clipped at 0 (cpp/include/mipt/small_rdm.hpp:172-277)
bool entropy(const std::vector<cplx> &rho, int d, double &s)
{
    std::vector<double> w;
    if (!hermitian_eigenvalues(rho, d, w)) return false;
    s = 0.0;
    for (double l : w)
    {
        if (!std::isfinite(l) || l < -2e-5) return false;
        if (l > 0.0) s -= l * std::log(l);
    }
    return true;
}

double purity(const std::vector<cplx> &rho, int d)
{
    double sq = 0.0, tr = 0.0;
    for (int r = 0; r < d; ++r)
    {
        tr += rho[r * d + r].real();
        for (int c = 0; c < d; ++c) sq += std::norm(rho[r * d + c]);
    }
    return sq / (tr * tr);
}

// 0.5 (|| rho^{T_j} ||_1 - 1), transpose of reference j (qubit convention)
bool cut_negativity(const std::vector<cplx> &rho, int k, int j, double &neg)
{
    const int D = 1 << k, m = 1 << j;
    std::vector<cplx> pt(static_cast<std::size_t>(D) * D);
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) pt[r * D + c] = rho[((r & ~m) | (c & m)) * D + ((c & ~m) | (r & m))];
    std::vector<double> w;
    if (!hermitian_eigenvalues(pt, D, w)) return false;
    double norm = 0.0;
    for (double l : w) norm += std::fabs(l);
    neg = 0.5 * (norm - 1.0);
    if (neg < 0.0 && neg > -1e-10) neg = 0.0;
    return std::isfinite(neg);
}
} // namespace

extern "C" int mipt_hip_probe_information(const double *rho_ri, int k, double *out)
{
    using mipt_hip::set_error;
    if (!rho_ri || !out || k < 1 || k > 4)
    {
        set_error("probe_information: bad arguments (k must be 1..4)");
        return mipt_hip::E_ARG;
    }
    const int D = 1 << k;
    std::vector<cplx> rho(static_cast<std::size_t>(D) * D);
    double tr = 0.0;
    for (int i = 0; i < D * D; ++i) rho[i] = cplx(rho_ri[2 * i], rho_ri[2 * i + 1]);
    for (int r = 0; r < D; ++r) tr += rho[r * D + r].real();
    if (!(tr > 0.0) || !std::isfinite(tr))
    {
        set_error("probe_information: invalid trace");
        return mipt_hip::E_NORM;
    }
    for (int i = 0; i < 11; ++i) out[i] = 0.0;
    // entropies and reduced matrices of every non-empty subset
    std::vector<double> S(static_cast<std::size_t>(D), 0.0);
    std::vector<std::vector<cplx>> sub(static_cast<std::size_t>(D));
    std::vector<int> dim(static_cast<std::size_t>(D), 0);
    for (unsigned mask = 1; mask < static_cast<unsigned>(D); ++mask)
    {
        sub[mask] = partial_trace(rho, k, mask, dim[mask]);
        if (!entropy(sub[mask], dim[mask], S[mask]))
        {
            set_error("probe_information: eigensolve failed or matrix not positive semidefinite");
            return mipt_hip::E_EIG;
        }
    }
    const unsigned all = static_cast<unsigned>(D) - 1u;
    if (k == 1)
    {
        out[0] = S[1];
        return mipt_hip::OK;
    }
    if (k == 2)
    {
        double mi = S[1] + S[2] - S[3];
        if (mi < 0.0 && mi > -2e-6) mi = 0.0;
        double neg = 0.0;
        if (!std::isfinite(mi) || mi < -2e-6 || !cut_negativity(sub[all], 2, 0, neg))
        {
            set_error("probe_information: invalid two-reference information");
            return mipt_hip::E_EIG;
        }
        out[0] = S[3];
        out[1] = std::max(0.0, mi);
        out[4] = std::max(0.0, neg);
        out[5] = std::log1p(2.0 * out[4]);
        return mipt_hip::OK;
    }
    if (k == 3)
    {
        const double iab = S[1] + S[2] - S[3], iac = S[1] + S[4] - S[5], ibc = S[2] + S[4] - S[6];
        out[6] = S[1] + S[2] + S[4] - S[3] - S[5] - S[6] + S[7];
        out[7] = (iab + iac + ibc) / 3.0;
        double best = INFINITY;
        for (int j = 0; j < 3; ++j)
        {
            double neg = 0.0;
            if (!cut_negativity(sub[all], 3, j, neg))
            {
                set_error("probe_information: cut negativity failed");
                return mipt_hip::E_EIG;
            }
            best = std::min(best, neg);
        }
        out[8] = std::max(0.0, best);
        out[9] = purity(sub[all], 8);
        out[10] = (purity(sub[1], 2) + purity(sub[2], 2) + purity(sub[4], 2)) / 3.0;
        return mipt_hip::OK;
    }
    double i2 = 0.0, i3 = 0.0, i4 = 0.0;
    int pairs = 0, triples = 0;
    for (unsigned mask = 1; mask <= all; ++mask)
    {
        const int n = __builtin_popcount(mask);
        i4 += (n % 2 ? 1.0 : -1.0) * S[mask];
        if (n == 2)
        {
            double v = -S[mask];
            for (int j = 0; j < 4; ++j)
                if ((mask >> j) & 1u) v += S[1u << j];
            i2 += v;
            ++pairs;
        }
        else if (n == 3)
        {
            double v = S[mask];
            for (unsigned s2 = mask; s2; s2 = (s2 - 1) & mask)
            {
                const int m = __builtin_popcount(s2);
                if (m == 1) v += S[s2];
                else if (m == 2) v -= S[s2];
            }
            i3 += v;
            ++triples;
        }
    }
    out[0] = S[all];
    out[1] = i2 / pairs;
    out[2] = i3 / triples;
    out[3] = i4;
    return mipt_hip::OK;
}
