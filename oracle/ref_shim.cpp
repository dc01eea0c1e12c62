// Thin extern "C" wrapper over the reference's own host-only headers, compiled
// straight from the reference tree (see oracle/Makefile) into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle against the
// reference code where that code builds without CUDA, CUDA-Q or LAPACK.
//
// Wrapped reference code (paths relative to the reference repository root):
//   cpp/include/mipt/free_energy_measure.hpp:19-95  measure_collapse_host
//   cpp/include/mipt/tmi/linalg.hpp:125-141, 262-352, 524-583, 1237-1256
//       (entropy cut-offs, Jacobi fallback, entropy_subsystem_rdm_from_state;
//        no LAPACK is installed, so the Jacobi path is the one exercised)
//   cpp/include/mipt/util/spectral.hpp:40-160  hermitian_eigenvalues<N>
//   cpp/include/mipt/types.hpp:23-38, 262-273  linspace, splitmix64, trajectory_seed
//   cpp/include/mipt/cusv/matrix.hpp:214-235  tensor_product
#include "mipt/cusv/matrix.hpp"
#include "mipt/free_energy_measure.hpp"
#include "mipt/tmi/linalg.hpp"
#include "mipt/types.hpp"
#include "mipt/util/spectral.hpp"

#include <complex>
#include <cstdint>
#include <vector>

template <std::size_t N>
static int eig_n(const double *a_ri, double *w)
{
    std::array<std::complex<double>, N * N> a{};
    for (std::size_t i = 0; i < N * N; ++i) a[i] = {a_ri[2 * i], a_ri[2 * i + 1]};
    std::array<double, N> ev{};
    if (!mipt::util::hermitian_eigenvalues<N>(a, ev)) return 1;
    for (std::size_t i = 0; i < N; ++i) w[i] = ev[i];
    return 0;
}

extern "C" {

int ref_measure_collapse_host(double *psi_ri, int n, int q, double u, int *outcome, double *prob, double *surprisal)
{
    try
    {
        auto r = mipt::free_energy::measure_collapse_host(reinterpret_cast<std::complex<double> *>(psi_ri), n, q, u);
        *outcome = r.outcome;
        *prob = r.conditional_probability;
        *surprisal = r.surprisal;
        return 0;
    }
    catch (...)
    {
        return -1;
    }
}

// Entropy (bits) of the kept set of a complex128 state: manual RDM then
// entropy_hermitian_inplace (order 1 = von Neumann, 2 = Renyi-2).
double ref_entropy_subsystem_rdm(const double *psi_ri, int n, const int *kept, int count, int order)
{
    std::vector<int> modes(kept, kept + count);
    mipt::tmi::StateEntropyWorkspace ws;
    return mipt::tmi::entropy_subsystem_rdm_from_state(
        reinterpret_cast<const std::complex<double> *>(psi_ri), n, modes, false, ws,
        order == 2 ? mipt::tmi::EntropyOrder::Renyi2 : mipt::tmi::EntropyOrder::VonNeumann);
}

// entropy_hermitian_inplace on a row-major interleaved matrix.
double ref_entropy_hermitian(const double *a_ri, int n, int order)
{
    std::vector<mipt::tmi::C64> a(static_cast<std::size_t>(n) * n);
    for (std::size_t i = 0; i < a.size(); ++i) a[i] = mipt::tmi::C64(a_ri[2 * i], a_ri[2 * i + 1]);
    mipt::tmi::EntropyWorkspace ws;
    return mipt::tmi::entropy_hermitian_inplace(
        a, static_cast<std::size_t>(n), ws,
        order == 2 ? mipt::tmi::EntropyOrder::Renyi2 : mipt::tmi::EntropyOrder::VonNeumann);
}

double ref_entropy_from_eigenvalue(double lambda) { return mipt::tmi::entropy_from_eigenvalue(lambda); }
double ref_finish_entropy(double s) { return mipt::tmi::finish_entropy(s); }

int ref_hermitian_eigenvalues(const double *a_ri, int n, double *w)
{
    switch (n)
    {
    case 2: return eig_n<2>(a_ri, w);
    case 4: return eig_n<4>(a_ri, w);
    case 8: return eig_n<8>(a_ri, w);
    case 16: return eig_n<16>(a_ri, w);
    default: return -1;
    }
}

std::uint64_t ref_splitmix64(std::uint64_t v) { return mipt::seeding::splitmix64(v); }
std::uint64_t ref_trajectory_seed(std::uint64_t master, std::uint64_t index)
{
    return mipt::seeding::trajectory_seed(master, index);
}

void ref_linspace(double first, double last, int count, double *out)
{
    const auto v = mipt::linspace(first, last, count);
    for (std::size_t i = 0; i < v.size(); ++i) out[i] = v[i];
}

// tensor_product(low, high) of two dense matrices given as interleaved
// row-major complex arrays with k_low / k_high targets.
void ref_tensor_product(const double *low_ri, int k_low, const double *high_ri, int k_high, double *out_ri)
{
    mipt::cusv::DenseMatrix lo(k_low), hi(k_high);
    for (std::size_t i = 0; i < lo.values.size(); ++i) lo.values[i] = {low_ri[2 * i], low_ri[2 * i + 1]};
    for (std::size_t i = 0; i < hi.values.size(); ++i) hi.values[i] = {high_ri[2 * i], high_ri[2 * i + 1]};
    const auto out = mipt::cusv::tensor_product(lo, hi);
    for (std::size_t i = 0; i < out.values.size(); ++i)
    {
        out_ri[2 * i] = out.values[i].real();
        out_ri[2 * i + 1] = out.values[i].imag();
    }
}

} // extern "C"
