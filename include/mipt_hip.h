/*
 * libmipt_hip: MI355X (gfx950) qubit MIPT trajectory engine, C ABI.
 *
 * Conventions shared by every entry point:
 *   - Return 0 on success.  Negative values are argument/numeric errors
 *     (-1 invalid argument, -2 invalid branch norms or zero state,
 *      -3 invalid conditional probability, -4 invalid branch norm,
 *      -5 unsupported request, -6 out of memory); positive values are
 *     hipError_t codes, 100000+status rocSOLVER/rocBLAS failures and
 *     102000+info an FP64 eigensolve that did not converge.
 *     mipt_hip_last_error() describes the last failure of the calling thread.
 *   - State vectors are device pointers to interleaved complex numbers
 *     (complex64 for *_f32, complex128 for *_f64); qubit q is index bit q.
 *   - Host buffers are caller-owned.  Calls are synchronous unless stated.
 *   - Entropies of TMI/half-chain subsystems are in bits; probe entropies in
 *     nats (the units of the reference).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   cpp/include/mipt/cuda/tmi_cuda_svd.hpp:20-60     entropy / Renyi-2 C ABI
 *   cpp/include/mipt/cuda/tmi_cuda_rdm.hpp:20-33     subsystem RDM export
 *   cpp/include/mipt/cuda/free_energy_measure.hpp:13-29  sequential measurement
 *   cpp/include/mipt/cuda/ancilla_rdm.hpp:19-29      probe rho1
 *   cpp/include/mipt/cusv/engine.hpp:100-582         per-trajectory engine
 *   cpp/include/mipt/circuit.hpp:240-282, 409-443    Haar trajectory loop
 */
#ifndef MIPT_HIP_H
#define MIPT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *mipt_hip_last_error(void);
/* Library version string. */
const char *mipt_hip_version(void);

/* Measurement helper (no counterpart in the reference): copies `bytes`
 * (a multiple of 16) device to device with a plain streaming kernel, so that
 * hardware traffic counters can be calibrated against a known byte count. */
int mipt_hip_stream_copy(void *dst, const void *src, uint64_t bytes, void *hip_stream);

/* ------------------------------------------------------------------ drop-ins */

/* Replaces mipt_cuda_measure_collapse_f32/_f64
 * (cpp/include/mipt/cuda/free_energy_measure.hpp:13-27): P(1) from FP64 branch
 * norms of the whole state, outcome = 1 iff uniform_01 < P(1), collapse by
 * 1/sqrt(raw branch norm).  realized_probability = conditional P of the outcome. */
int mipt_hip_measure_collapse_f32(void *device_state_vector, int n_qubits, int measured_qubit, double uniform_01,
                                  int *outcome, double *realized_probability);
int mipt_hip_measure_collapse_f64(void *device_state_vector, int n_qubits, int measured_qubit, double uniform_01,
                                  int *outcome, double *realized_probability);
/* Replaces mipt_cuda_measure_collapse_last_error (free_energy_measure.hpp:29). */
const char *mipt_hip_measure_collapse_last_error(void);
/* No counterpart in the reference: mipt_hip_measure_collapse_f32 for callers
 * that cannot allocate device memory (tight serving loops, harnesses that
 * forbid hipMalloc).  Same kernels, sums, outcome, probability and collapse;
 * its working memory is a fixed 1 MiB device array per GPU, so states of at
 * most 29 qubits (larger n returns -5 with a message).  Calls that overlap on
 * one GPU race on that array. */
int mipt_hip_measure_collapse_f32_noalloc(void *device_state_vector, int n_qubits, int measured_qubit,
                                          double uniform_01, int *outcome, double *realized_probability);

/* Replace mipt_cuda_entropy_svd_f32/_f64 (cpp/include/mipt/cuda/tmi_cuda_svd.hpp:20-34):
 * von Neumann entropy (bits) of the kept modes (little-endian order) of a pure
 * device state.  RDM of the complex64 state by a Gram on the matrix cores
 * (FP64 MFMA by default, FP32 MFMA with FP64 promotion when selected below),
 * FP64 for _f64; FP64 Hermitian eigenvalues in both.  fermion_trace must be 0
 * (qubit trace). */
int mipt_hip_entropy_svd_f32(const void *device_state_vector, int n_qubits, const int *host_kept_modes,
                             int kept_count, int fermion_trace, double *entropy_out);
int mipt_hip_entropy_svd_f64(const void *device_state_vector, int n_qubits, const int *host_kept_modes,
                             int kept_count, int fermion_trace, double *entropy_out);

/* Gram arithmetic of every complex64 RDM (no counterpart in the reference,
 * whose Cherk is FP32): 64 = FP64 MFMA (default), 32 = FP32 MFMA with FP64
 * promotion every 16-deep K step.  Process-wide; overrides
 * MIPT_HIP_GRAM_PRECISION.  -1 for any other value. */
int mipt_hip_set_gram_precision(int bits);

/* Eigenvalue path of every entropy (no counterpart in the reference, whose
 * device entropy uses cuSOLVER heevd/gesvd, tmi_cuda_svd.cu:663-798):
 * 0 = the library's FP64 kernels (default): blocked Householder reduction to
 * a real tridiagonal and bisection on Sturm counts, one workgroup per matrix,
 * for batches of RDMs up to 1024 x 1024 and single RDMs up to 512 x 512,
 * rocSOLVER zheevd beyond; 1 = rocSOLVER zheevd for every size.
 * Process-wide.  -1 for any other value. */
int mipt_hip_set_eigensolver(int which);

/* Replaces entropy_hermitian_inplace (cpp/include/mipt/tmi/linalg.hpp:524-583)
 * for many matrices at once: von Neumann entropies (bits; lambda <= 1e-15
 * dropped, |S| < 1e-12 -> 0 as linalg.hpp:125-141) of `count` Hermitian m x m
 * host matrices, row-major interleaved FP64 like the RDM export
 * (rho_ri[2 (i m^2 + r m + c) + {0,1}]), on the library's batched FP64
 * eigensolver; m <= 1024.  eigenvalues (ascending, count * m), may be NULL. */
int mipt_hip_entropy_hermitian_batched(const double *rho_ri, int m, int count, double *entropy_bits,
                                       double *eigenvalues);

/* Measurement helpers (no counterpart in the reference): time every Gram
 * kernel of the complex64 RDM path with HIP events on its stream (enable 1;
 * 0 stops; either clears earlier timings), then the average kernel duration
 * in ms, how many were timed and their average algorithmic FLOPs
 * (4 m (m + 1) K for the lower triangle of an m x K complex M M^H). */
int mipt_hip_set_gram_profiling(int enable);
int mipt_hip_gram_stats(double *avg_ms, int *count, double *flops_avg);

/* Replace mipt_cuda_renyi2_f32/_f64 (tmi_cuda_svd.hpp:36-58): S_2 = -log2 Tr rho^2 (bits). */
int mipt_hip_renyi2_f32(const void *device_state_vector, int n_qubits, const int *host_kept_modes, int kept_count,
                        int fermion_trace, double *entropy_out);
int mipt_hip_renyi2_f64(const void *device_state_vector, int n_qubits, const int *host_kept_modes, int kept_count,
                        int fermion_trace, double *entropy_out);

/* Replace mipt_cuda_rdm_subsystem_f32/_f64 (cpp/include/mipt/cuda/tmi_cuda_rdm.hpp:20-33):
 * rho[2*(r*D+c)+{0,1}] = Re/Im rho[r,c], D = 2^kept_count, kept_count <= 12. */
int mipt_hip_rdm_subsystem_f32(const void *device_state_vector, int n_qubits, const int *host_kept_modes,
                               int kept_count, int fermion_trace, double *host_rho_ri_row_major);
int mipt_hip_rdm_subsystem_f64(const void *device_state_vector, int n_qubits, const int *host_kept_modes,
                               int kept_count, int fermion_trace, double *host_rho_ri_row_major);

/* Replace mipt_cuda_rho1_complex_f32/_f64 (cpp/include/mipt/cuda/ancilla_rdm.hpp:19-29):
 * host_rho_ri = {rho00, rho11, Re rho01, Im rho01}, FP64 accumulation. */
int mipt_hip_rho1_complex_f32(const void *device_state_vector, int n_qubits, int retained_qubit,
                              double *host_rho_ri);
int mipt_hip_rho1_complex_f64(const void *device_state_vector, int n_qubits, int retained_qubit,
                              double *host_rho_ri);

/* TMI I3 = S_A+S_B+S_C+S_D-S_AB-S_AC-S_BC (bits) of cyclic blocks of `block`
 * sites starting at `cycle` (cpp/include/mipt/tmi/compute.hpp:89-113, 186-227);
 * D is the rest of the ring.  S(L/2) of kept {0..n/2-1}
 * (cpp/include/mipt/free_energy.hpp:274-287) is the entropy entry point above. */
int mipt_hip_tmi_f32(const void *device_state_vector, int n_qubits, int chain_sites, int block, int cycle,
                     double *tmi_out, double *terms_out /* 7: A,B,C,D,AB,AC,BC; may be NULL */);
/* The same TMI of a complex128 device state, FP64 Gram. */
int mipt_hip_tmi_f64(const void *device_state_vector, int n_qubits, int chain_sites, int block, int cycle,
                     double *tmi_out, double *terms_out /* 7: A,B,C,D,AB,AC,BC; may be NULL */);

/* ---------------------------------------------------------- plain engine */
/* One trajectory at a time, one launch per blocked op, a host round trip per
 * measured layer: the semantics and structure of cusv::Engine
 * (cpp/include/mipt/cusv/engine.hpp:100-582). */
typedef struct mipt_hip_engine mipt_hip_engine;

/* stream may be NULL (the null stream). */
int mipt_hip_engine_create(int use_fp64, void *hip_stream, mipt_hip_engine **out);
int mipt_hip_engine_destroy(mipt_hip_engine *engine);
/* Operate on a caller-owned device buffer (engine.hpp:131-140). */
int mipt_hip_engine_adopt(mipt_hip_engine *engine, void *device_state_vector, int n_qubits);
/* Engine-owned |0...0> of n qubits (engine.hpp:143-160). */
int mipt_hip_engine_reset(mipt_hip_engine *engine, int n_qubits);
void *mipt_hip_engine_data(mipt_hip_engine *engine);
/* Dense 2^k x 2^k row-major matrix (interleaved complex128, narrowed to the
 * state precision) on k <= 5 targets; bit j of the matrix index = targets[j]
 * (engine.hpp:236-277). */
int mipt_hip_engine_apply_matrix(mipt_hip_engine *engine, const int *targets, int k, const double *matrix_ri);
/* A layer of two-qubit gates (q0 = bit 0 of each 4x4) fused into Kronecker
 * blocks of at most block_targets qubits (cpp/include/mipt/cusv/gates.hpp:295-333). */
int mipt_hip_engine_apply_bonds(mipt_hip_engine *engine, const int *q0, const int *q1, const double *gates_ri,
                                int count, int block_targets);
/* Measures `sites` (ascending) and collapses.  mode 0: sequential per-site
 * draws uniforms[j] (outcome 1 iff u < P(1)); mode 1: joint draw per chunk of
 * chunk_bits sites with uniforms[chunk] (engine.hpp:293-368).  p1 (mode 0,
 * may be NULL) receives each conditional P(1); surprisal = -log P(outcomes). */
int mipt_hip_engine_measure_layer(mipt_hip_engine *engine, const int *sites, int count, const double *uniforms,
                                  int mode, int chunk_bits, int *outcomes, double *p1, double *surprisal);
/* Replaces the per-trajectory loop Circuit1D::try_cusv
 * (cpp/include/mipt/circuit.hpp:240-282) on this engine in one call: reset to
 * |0...0>, then per layer the bonds of geometry (n, closed, grid_x: 0 = chain,
 * else grid columns) fused into Kronecker blocks of at most block_targets
 * qubits and applied, then the flagged sites measured as
 * mipt_hip_engine_measure_layer does (mode, chunk_bits), with the per-layer
 * histogram copied to the host.  Every layer's blocked matrices are staged to
 * the device in one copy before the first layer.  gates: interleaved complex128
 * [total bonds][4][4] in layer order; flags, uniforms: [layers][n] (mode 1:
 * uniforms[t][chunk]); outcomes [layers][n] out, -1 where unmeasured. */
int mipt_hip_engine_run_trajectory(mipt_hip_engine *engine, int n, int layers, int closed, int grid_x,
                                   const double *gates_ri, const uint8_t *flags, const double *uniforms,
                                   int block_targets, int mode, int chunk_bits, int *outcomes);
int mipt_hip_engine_to_host(mipt_hip_engine *engine, void *host_buffer);
int mipt_hip_engine_set_from_host(mipt_hip_engine *engine, const void *host_buffer);
int mipt_hip_engine_squared_norm(mipt_hip_engine *engine, double *out);

/* ---------------------------------------------------------- batched engine */
/* Thousands of trajectories (mixed chain lengths) evolved on device-resident
 * states with fused tile sweeps; no host round trip inside the evolution. */
typedef struct mipt_hip_batch mipt_hip_batch;

typedef struct mipt_hip_traj_spec
{
    int L;              /* chain sites (gates act on 0..L-1) */
    int layers;         /* brickwork layers (2 x periods) */
    double p;           /* measurement probability per site per layer */
    uint64_t seed;      /* counter-based stream key */
    int closed;         /* 1: ring (wrap bond on odd layers), 0: open chain */
    int gate_family;    /* 1 = Haar U(4) (cpp/include/mipt/types.hpp:155-186) */
    int probe_site;     /* -1: none; else a probe qubit at index L is attached to this site: Bell-paired
                           at t=0 when p_zero_layers = 0, else reset(site), h(site), cx(site, probe) at
                           the start of layer p_zero_layers (cpp/include/mipt/probed/kernels.hpp:68-77) */
    int p_zero_layers;  /* the first p_zero_layers layers are unmeasured (scrambling) */
    int shared_gates;   /* -1: own gates; else use trajectory shared_gates's gate stream */
    int grid_x;         /* 0: chain; else an L = grid_x * (L / grid_x) grid, site r*grid_x + c, with four
                           bond colours per time step (layer mod 4): horizontal from even / odd columns
                           (odd colour closes each row when closed and grid_x is even), vertical from
                           staggered even / odd rows (cpp/include/mipt/circuits/mms.hpp:159-330) */
    int probe_count;    /* 0: one probe iff probe_site >= 0 (above).  1..4 (probe_site then -1): reference j
                           is qubit L + j, paired with the distinct site probe_sites[j]; all are attached in
                           order j = 0, 1, ... at the start of layer attach_layer (reset, h, cx each; the
                           reset's uniform is keyed by (attach_layer, site); the previous layer's measurements
                           are applied first), or Bell-paired at t = 0 when attach_layer = 0
                           (cpp/include/mipt/probed/kernels.hpp:33-66) */
    int probe_sites[4];
    int attach_layer;   /* with probe_count >= 1: the layer the references attach at (-1: p_zero_layers);
                           independent of the unmeasured prefix, e.g. two references attached at t0 = 2N after
                           monitored layers (cpp/include/mipt/probed/runner.hpp:258-307) */
    int attach_stagger; /* with probe_count >= 2: reference j attaches at attach_layer + j * attach_stagger
                           (0: all together); a site may then serve several references, as the staggered
                           insertion of cpp/include/mipt/probed/runner.hpp:230-256 does */
    int fp64_check;     /* 1: mipt_hip_run_batch states this trajectory's errors against a complex128 replay
                           of the same FP64 gates, flags and uniforms (traj_result amp_err, s_half_err,
                           tmi_err); 0: those fields are NaN */
} mipt_hip_traj_spec;

typedef struct mipt_hip_batch_config
{
    int tile_bits;      /* 0 = the autotuner's choice (mipt_hip_batch_autotune; 13 before it runs); else 12..14 */
    int max_hist_bits;  /* measured sites per fused histogram chunk (<= 10); 0 = default */
    int threads;        /* with tile_bits = 13: 128 selects 64-amplitude registers per thread; else 0 */
    int bond_arith;     /* bond arithmetic of the tile sweeps: 0 = the autotuner's choice (matrix cores
                           before it runs); 1 = VALU only; 2 = VALU and matrix cores; 3 = matrix cores
                           only (v_mfma_f32_4x4x1_16b_f32) */
    int plan;           /* layer plan variant 0..3 (qubit-block assignment, see mipt_hip_batch_autotune);
                           replaced by the autotuner's choice once it has run */
} mipt_hip_batch_config;

/* The batch owns every device buffer its calls use (states, draws, histogram
 * partials, and the entropy workspace of its readouts with its rocSOLVER
 * handle); destroy synchronises its streams and releases all of it. */
int mipt_hip_batch_create(const mipt_hip_batch_config *config, void *hip_stream, mipt_hip_batch **out);
int mipt_hip_batch_destroy(mipt_hip_batch *batch);
/* Allocates device-resident states and streams for `count` trajectories. */
int mipt_hip_batch_setup(mipt_hip_batch *batch, const mipt_hip_traj_spec *specs, int count);
/* Draws every gate, flag and uniform on device from the counter-based stream
 * (Philox4x32-10 keyed by the seed; asynchronous on the batch stream). */
int mipt_hip_batch_generate(mipt_hip_batch *batch);
/* Replay: explicit unitaries (complex128 [bonds][4][4] interleaved, bond order),
 * flags ([layers][L] 0/1) and per-site uniforms ([layers][L]) for one trajectory. */
int mipt_hip_batch_set_layers(mipt_hip_batch *batch, int traj, const double *gates_ri, const uint8_t *flags,
                              const double *uniforms);
/* Initial state (complex64 interleaved, 2^(L+probe) amplitudes); default |0...0>
 * (or the probe Bell pair). */
int mipt_hip_batch_set_state(mipt_hip_batch *batch, int traj, const float *psi_ri);
int mipt_hip_batch_get_state(mipt_hip_batch *batch, int traj, float *psi_ri);
/* Device pointer of a trajectory's state (valid until destroy). */
void *mipt_hip_batch_state_ptr(mipt_hip_batch *batch, int traj);
/* Evolves every trajectory through all of its layers (asynchronous). */
int mipt_hip_batch_evolve(mipt_hip_batch *batch);
/* Evolves layers [layer_begin, layer_end) of all trajectories (asynchronous). */
int mipt_hip_batch_evolve_range(mipt_hip_batch *batch, int layer_begin, int layer_end);
int mipt_hip_batch_synchronize(mipt_hip_batch *batch);
/* Autotuner: times probe_layers layers of this batch (its own draws and
 * shapes) under each tile configuration -- 0: 2^13-amplitude tiles, 256
 * threads, bond arithmetic on the VALU; 1: the same split between the VALU
 * and the matrix cores (v_mfma_f32_4x4x1_16b_f32); 2: the same all on the
 * matrix cores; 3: 2^14 tiles, 256 threads; 4: 2^13 tiles, 128 threads;
 * 5: 2^12 tiles, 64 threads (3-5 on the VALU) -- and keeps the fastest for
 * the tile-swept trajectories of a batch created with tile_bits = 0.  The
 * batch is re-armed before and after (draws kept).  Then, on the winning
 * configuration, the qubit-block assignment: layer plan variants 1..3 (bit 0:
 * a bond local to several windows goes where it lands on a free register-slot
 * pair, needing no LDS exchange, before counts are balanced; bit 1: three or
 * more chain windows spread their overlaps evenly) against the default plan 0.
 * best_index (may be NULL) receives the configuration (-1: nothing is
 * tile-swept), best_plan (may be NULL) the plan variant, ms (may be NULL,
 * 9 doubles) each configuration's time, then plan variants 1..3 (0 where a
 * candidate does not apply). */
int mipt_hip_batch_autotune(mipt_hip_batch *batch, int probe_layers, int *best_index, int *best_plan, double *ms);
/* 1: full evolutions run as one captured HIP graph (both streams), captured
 * on the second full evolve after the draws or configuration last changed
 * and replayed afterwards; 0 (default): eager launches. */
int mipt_hip_batch_set_graph(mipt_hip_batch *batch, int enable);
/* outcomes [layers][L]: -1 unmeasured, else 0/1; flagged [layers][L]: 1 where
 * |u - P(1)| < 1e-6; p1 [layers][L] conditional P(1) (may be NULL). */
int mipt_hip_batch_outcomes(mipt_hip_batch *batch, int traj, int *outcomes, uint8_t *flagged, double *p1,
                            double *surprisal);
/* Per-trajectory observables: S(L/2) in bits, TMI (bits, cycle 0, block
 * tmi_block or L/4 if 0), probe entropy (nats) and Bloch vector (NaN when
 * the trajectory has no probe). */
int mipt_hip_batch_observables(mipt_hip_batch *batch, int traj, int tmi_block, double *s_half, double *tmi,
                               double *probe_entropy, double *bloch3);
/* TMI (bits) of cyclic blocks of `block` sites (0: L/4) starting at site
 * `cycle`, from von Neumann (order 1) or Renyi-2 (order 2) entropies; the
 * all_cycles scan of cpp/include/mipt/tmi/compute.hpp:115-118 calls it for
 * cycle = 0..L/4-1. */
int mipt_hip_batch_tmi(mipt_hip_batch *batch, int traj, int block, int cycle, int order, double *tmi);
/* S(L/2) (s_half[count], may be NULL) and von Neumann TMI for cycles
 * 0..tmi_cycles-1 (tmi[count][tmi_cycles]) of every trajectory, bits: the
 * RDMs of one size are solved together (rocSOLVER strided-batched zheevd),
 * which at L=20 is about 8x faster than one entropy call at a time. */
int mipt_hip_batch_observables_all(mipt_hip_batch *batch, int tmi_block, int tmi_cycles, double *s_half,
                                   double *tmi);
/* Translation-averaged entropies of cyclic blocks, the per-trajectory
 * quantity of cpp/include/mipt/entropy.hpp:521-577 (entropy.exe): for block
 * sizes l = min_kept..max_kept, the mean over the L starts i of the von
 * Neumann S1 (only for l <= max_s1_kept, else NaN) and Renyi-2 S2 =
 * -ln Tr(rho^2) of sites {i, ..., i+l-1 mod L}, in nats; outputs
 * [count][max_kept - min_kept + 1] (s1_nats may be NULL). */
int mipt_hip_batch_translation_entropies(mipt_hip_batch *batch, int min_kept, int max_kept, int max_s1_kept,
                                         double *s1_nats, double *s2_nats);
/* Bulk readout for learning from the simulator: the measurement records of
 * every trajectory for layers [layer_begin, layer_end) as int8
 * [count][layers][L] (-1 unmeasured, else the outcome; every trajectory must
 * have the same L), and the probe Bloch vectors [count][3]
 * ((2 Re rho01, -2 Im rho01, rho00 - rho11), cpp/include/mipt/ancilla_mi.hpp:
 * 24-52; NaN without a probe). */
int mipt_hip_batch_records(mipt_hip_batch *batch, int layer_begin, int layer_end, int8_t *host_records);
int mipt_hip_batch_probe_bloch(mipt_hip_batch *batch, double *host_bloch);
/* Joint density matrix of every trajectory's references (every trajectory needs the same count k >= 1):
 * host [count][2^k][2^k][re, im] FP64, trace 1, reduced-basis bit j = reference j (qubit L + j); the
 * multi-reference matrices of cpp/include/mipt/ancilla_rdm.hpp (rho2 / rho3 / rho4 of ancilla_rdm.cu). */
int mipt_hip_batch_probe_rdm(mipt_hip_batch *batch, double *host_rho);
/* Information measures (nats) of a k-reference density matrix (k = 1..4, [2^k][2^k][re, im]), in the
 * field order of the reference's probe sample (cpp/include/mipt/probed/sample.hpp:21-33):
 * out[0] sq, [1] i2, [2] i3, [3] i4, [4] negativity, [5] log_negativity, [6] tmi, [7] average_mi,
 * [8] min_bipartite_negativity, [9] joint_purity, [10] mean_single_purity; filled as
 * cpp/include/mipt/probed/kernels.hpp:106-147 fills them for k references (host code, no GPU). */
int mipt_hip_probe_information(const double *rho_ri, int k, double *out);
/* Three-site reduced density matrices of trajectories [first, first+count):
 * subsystems [nsub][3] (qubit k of a subsystem = reduced-basis bit k; the
 * cyclic (i, i+1, i+2) list of cpp/include/mipt/rdm.hpp:92-106, or the
 * horizontal-then-vertical grid list of rdm.hpp:108-125), output
 * [count][nsub][8][8][re, im] FP64, trace-normalised -- the record payload of
 * the MIPTRHO2 file (cpp/include/mipt/density.hpp:10-48). */
int mipt_hip_batch_rho3(mipt_hip_batch *batch, int first, int count, const int *subsystems, int nsub,
                        double *host_rho_ri);
/* The stream the batch launches on. */
void *mipt_hip_batch_stream(mipt_hip_batch *batch);
/* Re-arms the in-kernel initial states and clears per-run records, so the
 * same trajectories can be run again (generate + evolve). */
int mipt_hip_batch_reset(mipt_hip_batch *batch);
/* Copies a trajectory's device-drawn unitaries (complex128 [bonds][4][4]),
 * flags ([layers][L]) and uniforms ([layers][L]) to host (any may be NULL). */
int mipt_hip_batch_export_layers(mipt_hip_batch *batch, int traj, double *gates_ri, uint8_t *flags,
                                 double *uniforms);
/* Layer fusion (no counterpart in the reference; process-wide, default on,
 * else MIPT_HIP_FUSE=0/1): the sweeps of consecutive layers alternate their
 * window order, and a trajectory that measures at most one site in a layer,
 * not bonded in the next layer's first sweep, runs that sweep in the same
 * pass over its state as this layer's last, the layer's collapse following
 * at the next layer's second sweep (2^13-amplitude tiles of 256 threads, no
 * reference attached during the evolution).  Results agree with the unfused
 * schedule to rounding (commuting bonds and projectors meet in another
 * order).  on = -1 goes back to
 * MIPT_HIP_FUSE / the default.  Status -1 for any other value. */
int mipt_hip_set_layer_fusion(int on);
/* Tile plan of one layer (no counterpart in the reference; for callers and
 * tests that check or inspect the sweep schedule on the host, no GPU needed):
 * the sweeps the batch runs for layer `layer` of an nq-qubit state whose
 * first L qubits form the chain (closed ring if `closed`, grid_x columns for
 * a grid, 0 for a chain) under tile_bits 12..14 / threads as in
 * mipt_hip_batch_config (tile_bits 13 with threads 128 or 256), layer plan
 * variant 0..3.  out receives ints: nbonds, then (q0, q1) of every bond of
 * the layer; npasses; per pass k, c, s, nstages; per stage use_lds, zpair,
 * then (bond, q0, q1) for bond[0] and bond[1] (bond -1: none), reg_u[6],
 * lane_u[6], wave_u[4].  n_out receives the count of ints; status -1 (and
 * nothing written) when out_len is smaller. */
int mipt_hip_layer_plan(int L, int nq, int closed, int grid_x, int layer, int tile_bits, int threads, int plan,
                        int *out, int out_len, int *n_out);
/* The sweep kernels compiled for a fixed stage structure (no counterpart in
 * the reference; for tests and callers that check which plans get them, no
 * GPU needed): out receives (zp, pat, pat2) per kernel.  zp 0..2: a sweep
 * whose qubit-0 form (the zpair of its last stage) is zp and whose structure
 * code is pat; zp 4..6: a fused launch, a form-0 sweep of code pat followed
 * by the next layer's first sweep of form zp - 4 and code pat2.  The code of
 * a plan of 1..3 stages: its stage count, plus per stage si the bits
 * 1 << (2 + 3 si) (the stage starts with an exchange), 1 << (3 + 3 si)
 * (bond[0] present), 1 << (4 + 3 si) (bond[1] present).  Used for 2^13-
 * amplitude tiles of 256 threads with matrix-core bonds; other sweeps run a
 * stage loop that reads the structure at run time.  n_out receives the count
 * of ints; status -1 (nothing written) when out_len is smaller. */
int mipt_hip_compiled_sweeps(int *out, int out_len, int *n_out);
/* Tiled groups in flight: 2 (default) alternates groups between the batch
 * stream and a second stream so one group's small kernels overlap the
 * other's sweeps; 1 runs every group on the batch stream, so a sweep's
 * event-timed duration is its own (per-kernel roofline measurement). */
int mipt_hip_batch_set_streams(mipt_hip_batch *batch, int streams);
/* Gate-sweep timing: every sample_every-th sweep kernel is bracketed by HIP
 * events on the batch stream (0 = off); resets the statistics. */
int mipt_hip_batch_set_profiling(mipt_hip_batch *batch, int sample_every);
/* Average duration (ms) of the timed sweeps, how many were timed, their mean
 * algorithmic bytes (one state read + write per layer, split over the layer's
 * sweeps), all sweeps launched and their total algorithmic bytes. */
int mipt_hip_batch_pass_stats(mipt_hip_batch *batch, double *avg_ms, int *timed, double *bytes_avg,
                              long *launches, double *bytes_total);

/* ------------------------------------------------------- learned decoder */
/* A small space-time convolutional network (no counterpart in the
 * reference) that reads a measurement record int8 [height=time][width=sites]
 * (-1 unmeasured, 0, 1) and predicts the probe qubit's conditional Bloch
 * vector.  Input channels: record == 0, record == 1, the probe's column
 * (x == width/2), time fraction (y+1)/height; four 3x3 convolutions to 64
 * channels (zero padding in time, periodic in sites) with ReLU, mean
 * pooling, 64 -> 64 (ReLU) -> 3.  BF16 weights/activations on the matrix
 * cores with FP32 accumulation, FP32 master weights and Adam (betas
 * 0.9/0.999, eps 1e-8).  Loss: batch mean of h(S(|out|) - S(|y|)) +
 * (|y| - out.y/|out|), |v| = sqrt(sum v^2 + 1e-12), S the qubit entropy
 * (nats) of Bloch length r, continued linearly past r = 1 - 1e-6, h the
 * smooth-L1 function of width delta (default 0.05; 0 = squared error).
 * Parameters are one FP32 vector: per conv layer w[9 taps][CI][64] then
 * b[64] (CI = 4, 64, 64, 64; tap = 3(dy+1) + (dx+1)), then fc1 w[64][64],
 * b[64], fc2 w[64][3], b[3]. */
typedef struct mipt_hip_decoder mipt_hip_decoder;
int mipt_hip_decoder_create(int height, int width, int batch_max, uint64_t seed, mipt_hip_decoder **out);
int mipt_hip_decoder_destroy(mipt_hip_decoder *decoder);
int mipt_hip_decoder_param_count(mipt_hip_decoder *decoder);
int mipt_hip_decoder_get_params(mipt_hip_decoder *decoder, float *host_params);
/* Also resets the optimiser state. */
int mipt_hip_decoder_set_params(mipt_hip_decoder *decoder, const float *host_params);
/* Uploads a dataset: records int8 [count][height][width], labels FP32 [count][3] (may be NULL). */
int mipt_hip_decoder_set_dataset(mipt_hip_decoder *decoder, const int8_t *host_records, const float *host_labels,
                                 int count);
/* Loss and gradient (FP32, parameter layout) of the batch of dataset rows host_idx[0..batch). */
int mipt_hip_decoder_gradients(mipt_hip_decoder *decoder, const int *host_idx, int batch, float *host_grad,
                               double *loss);
/* `steps` Adam steps, step s on rows host_idx[s*batch .. (s+1)*batch); mean batch loss (may be NULL). */
int mipt_hip_decoder_train(mipt_hip_decoder *decoder, const int *host_idx, int batch, int steps, float lr,
                           double *mean_loss);
/* Predicted Bloch vectors [count][3] of dataset rows [first, first + count). */
int mipt_hip_decoder_predict(mipt_hip_decoder *decoder, int first, int count, float *host_out);
/* Width delta of the smooth-L1 entropy term (>= 0; 0 = squared error). */
int mipt_hip_decoder_set_loss(mipt_hip_decoder *decoder, float huber_delta);
int mipt_hip_decoder_synchronize(mipt_hip_decoder *decoder);

/* ------------------------------------------------------------ whole runs */
typedef struct mipt_hip_traj_result
{
    double s_half;        /* S(L/2), bits */
    double probe_entropy; /* probe entropy, nats; NaN without a probe */
    double bloch[3];      /* probe Bloch vector; NaN without a probe */
    double norm_error;    /* | <psi|psi> - 1 | of the final complex64 state */
    int measured;         /* measured draws of the trajectory */
    int flagged_draws;    /* draws within 1e-6 of their Born threshold */
    /* Errors of this result, stated when the spec asks for fp64_check: the
     * trajectory is replayed on the plain engine in complex128 with its gates
     * drawn in FP64 by the same counter-based sampler (before the rounding to
     * complex64) and its flags and uniforms, and S and TMI of that state are
     * taken in FP64.  NaN when not asked for, for trajectories the plain
     * engine does not replay (probes, references, an unmeasured prefix,
     * shared gates), when a near-threshold draw came out differently in the
     * replay, and tmi_err for entropy order 2. */
    double amp_err;       /* max |psi - psi_128| over the amplitudes */
    double s_half_err;    /* |S(L/2) - S(L/2) of the replay|, bits */
    double tmi_err;       /* max over the cycles |TMI - TMI of the replay|, bits */
} mipt_hip_traj_result;

/* One call per batch of trajectories (the batched entry point a scan driver
 * needs in place of cpp/include/mipt/circuit.hpp:409-443 + tmi/compute.hpp:
 * 704-839 per trajectory): device draws, evolution, then per trajectory S(L/2),
 * TMI for cycles 0..tmi_cycles-1 (block 0 = L/4; entropy_order 1 = von
 * Neumann, 2 = Renyi-2; tmi_out[traj * tmi_cycles + cycle]), probe readout and
 * draw statistics into caller-owned host arrays. */
int mipt_hip_run_batch(const mipt_hip_traj_spec *specs, int count, int tmi_block, int tmi_cycles, int entropy_order,
                       double *tmi_out, mipt_hip_traj_result *results, void *hip_stream);

/* The online TMI scan of cpp/apps/sim_tmi.cpp:11-90 and
 * cpp/include/mipt/tmi/runner.hpp:88-409 on this engine: the same positional
 * arguments, CSV `p,tmi` and default output path; returns the exit code
 * (0, 1 on an error, 2 on too many arguments). */
int mipt_hip_sim_tmi_main(int argc, char **argv);

/* The three-site RDM record of cpp/apps/mipt.cpp:79-197, 309-400 on this
 * engine: [d] [n] [periods] [realizations] [resolution] [p_min] [p_max]
 * [circ_type]; writes info.csv and rho3.bin (d = 1) or rho3_2d.bin (d = 2) in
 * the MIPTRHO2 version-2 format of cpp/include/mipt/density.hpp:10-48; returns
 * the exit code. */
int mipt_hip_mipt_main(int argc, char **argv);

/* entropy.exe on this engine (cpp/apps/entropy.cpp:16-80): `N periods p
 * realizations circ_type output.csv`, translation-averaged S1/S2 (nats) per
 * block size with Welford statistics over realizations, the CSV schema of
 * cpp/include/mipt/entropy_resume.hpp:52-56; returns the exit code. */
int mipt_hip_entropy_main(int argc, char **argv);
/* mipt_probed.exe, modes 0 to 3 and two-reference 4 (cpp/apps/mipt_probed.cpp; probed/config.hpp:261-470,
 * runner.hpp:258-307, output.hpp:36-95, 386-476): [probes] [N] [realizations] [circ_type] [mode] [args] with
 * probes 1, 2 or 4, circ_type 1; the same reference sites, encoder schedule, readout times, output name and CSV
 * columns.  Returns 0, or 1 on an error (message on stderr). */
int mipt_hip_probed_main(int argc, char **argv);
/* free_energy.exe (cpp/apps/free_energy.cpp): [N] [p] [realizations] [circ_type] [init_state] with
 * circ_type 1; 4N equilibration + 24N production layers, cumulative measurement surprisal (nats) and
 * S(L/2) (bits, t <= MIPT_FREE_ENERGY_ENTROPY_MAX_T); the reference's samples and time-series CSVs
 * (free_energy_resume.hpp:79-88) under its generated names.  No resume: existing CSVs are refused
 * unless MIPT_FREE_ENERGY_RESUME=0.  Returns 0, 1 on an error, 2 on a usage error. */
int mipt_hip_free_energy_main(int argc, char **argv);
/* The initial state free_energy uses for a trajectory seed (host code): init_state 0 = product of
 * single-site states uniform on the Bloch sphere, 1 = normalised complex Gaussian vector
 * (cpp/include/mipt/free_energy.hpp:127-222); psi_ri [2^n][re, im]. */
int mipt_hip_free_energy_initial_state(int n, int init_state, uint64_t trajectory_seed, double *psi_ri);

#ifdef __cplusplus
}
#endif

#endif /* MIPT_HIP_H */
