// Batched trajectory engine: shared types of batch.hip and entropy.hip.
#pragma once

#include "common.hpp"
#include "mipt_hip.h"

#include <vector>

namespace mipt_hip
{

constexpr int MAX_HIST_BITS = 10;          // measured sites per fused histogram chunk
constexpr int HSTRIDE = 1 << MAX_HIST_BITS;  // doubles per tile partial
constexpr int AUTOTUNE_CANDIDATES = 6;    // tile shapes x bond arithmetic timed by the autotuner
constexpr int AUTOTUNE_PLANS = 4;         // layer plan variants (qubit-block assignment) timed on the winner
constexpr int SMALL_MAX_QUBITS = 14;       // whole state LDS-resident (128 KiB)

// Read-only per-trajectory description (device).
struct TrajDesc
{
    std::uint64_t state_off;  // complex64 element offset in the state arena
    std::uint64_t gate_off;   // gate (16 x complex64) offset in the gate arena
    std::uint32_t layer_off;  // offset into per-layer arrays (flags, outcomes, collapse)
    std::uint32_t unif_off;   // offset into per-(layer, site) arrays (uniforms, p1)
    int L, nq, layers, closed;
    int probe_site, p_zero_layers;  // probe_site: site of reference 0 (-1: no references)
    int gx, probes;                 // grid columns (0: chain); references (qubits L..L+probes-1)
    int psite[4];                   // site of reference j
    int attach, stagger;            // reference j attaches at layer attach + j * stagger (0: Bell pair at t = 0)
    double p;
    std::uint64_t seed;
};

// Pending collapse of one layer's measurements (phys bit q = qubit q).
struct Collapse
{
    std::uint64_t mask;
    std::uint64_t val;
    double scale;     // 1/sqrt(raw joint weight of the outcomes)
    double weight;    // raw joint weight of the decided chunks
};

// Device-resident arena of a batch.
struct DevArena
{
    float2 *states = nullptr;
    float2 *gates = nullptr;
    TrajDesc *desc = nullptr;
    std::uint32_t *flags = nullptr;      // [layer_off + t] site mask
    std::uint32_t *outcomes = nullptr;   // [layer_off + t] outcome mask
    std::uint32_t *flagged = nullptr;    // [layer_off + t] near-threshold mask
    double *uniforms = nullptr;          // [unif_off + t*L + s]
    double *p1 = nullptr;                // [unif_off + t*L + s]
    Collapse *coll = nullptr;            // [layer_off + t]
    double *surprisal = nullptr;         // [traj]
    int *status = nullptr;               // [traj]
    int *init = nullptr;                 // [traj] 1: build the initial state in-kernel at layer 0
    double *partials = nullptr;          // [group traj][tile][HSTRIDE]
    std::size_t partial_cap = 0;         // doubles
};

// Device workspace of the entropy paths (RDM buffers, eigensolver panels,
// the rocSOLVER handle).  A batch owns one, created with the batch's first
// readout and released with the batch; the context-free drop-in entry
// points of the reference's signatures use one per calling thread.
struct Workspace;
Workspace *workspace_create();
void workspace_destroy(Workspace *w);

// Entropy/observable helpers implemented in entropy.hip, used by the batch
// with its own workspace.
int entropy_bits_f32(Workspace &w, const float2 *psi, int n, const int *kept, int kept_count, hipStream_t stream,
                     double *out);
int tmi_bits_f32(Workspace &w, const float2 *psi, int n, int chain_sites, int block, int cycle, hipStream_t stream,
                 double *out, double *terms, int order = 1);
// Von Neumann entropies (bits) of many (state, kept set) items of one shape,
// with batched eigensolves; kept is [count][kept_count]
int entropy_many_bits_f32(Workspace &w, const float2 *const *psis, int n, const int *kept, int kept_count, int count,
                          hipStream_t stream, double *out);
// Tr(rho^2) of many (state, kept set) items of one shape
int purity_many_f32(Workspace &w, const float2 *const *psis, int n, const int *kept, int kept_count, int count,
                    hipStream_t stream, double *out);
// Renyi-2 entropy -log2 Tr(rho^2) of the kept set, bits
int renyi2_bits_f32(Workspace &w, const float2 *psi, int n, const int *kept, int kept_count, hipStream_t stream,
                    double *out);
int rho1_f32(Workspace &w, const float2 *psi, int n, int qubit, hipStream_t stream, double *out4);

// Batched FP64 Hermitian eigenvalues (eigen.hip): `count` m x m matrices at
// mats + i * stride (m <= EIG_MAX_N; the lower triangle of the column-major
// view is read, the matrices are overwritten), panel workspace wpanel of
// count * m * 32 elements, d and e of count * m doubles; entropies in bits
// to ent[count] and, when lam is not null, eigenvalues to lam[count * m].
// Launches on `stream` without synchronising.
constexpr int EIG_MAX_N = 1024;
int launch_hermitian_entropies(double2 *mats, int m, std::size_t stride, int count, double2 *wpanel, double *d,
                               double *e, double *ent, double *lam, hipStream_t stream);

} // namespace mipt_hip
