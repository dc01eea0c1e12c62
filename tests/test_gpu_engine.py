"""GPU parity of the HIP engine against the CPU FP64 oracle (replay mode).

Bar (stated per test): measurement outcomes bit-exact except draws within
1e-6 of the Born threshold (flagged); complex64 amplitudes within 1e-5 of the
complex128 oracle; FP64 paths within 1e-11.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FLAG_EPS = 1e-6


@pytest.fixture(scope="module")
def M():
    import mipt_amd
    mipt_amd.lib()
    return mipt_amd


class DevState:
    """A device state owned by a plain engine (the library's own HIP runtime;
    no second GPU runtime is loaded into the test process)."""

    def __init__(self, M, psi, fp64=False):
        n = int(np.log2(len(psi)))
        self.e = M.PlainEngine(fp64)
        self.e.reset(n)
        self.e.set_from_host(np.asarray(psi))

    def data_ptr(self):
        return self.e.data


@pytest.fixture(scope="module")
def gpu(M):
    """The library module; device buffers come from its own HIP runtime."""
    return M


def dev_state(M, psi, fp64=False):
    return DevState(M, psi, fp64)


def host_state(t, fp64=False):
    return t.e.to_host()


def check_outcomes(got, want, p1_oracle, unif, flags, label):
    """Outcomes equal the oracle's except draws within FLAG_EPS of P(1)."""
    m = flags == 1
    near = np.abs(unif - p1_oracle) < FLAG_EPS
    bad = m & (got != want) & ~near
    assert not bad.any(), f"{label}: outcome mismatch at {np.argwhere(bad)[:5]}"
    assert np.all((got >= 0) == m), f"{label}: measured-site pattern differs"


# ------------------------------------------------------------------ drop-ins

@pytest.mark.parametrize("fp64", [False, True])
def test_measure_collapse_kat(M, gpu, fp64):
    # tests/core_tests.cpp:120-157 through the device C ABI
    print(f"measure_collapse KAT fp64={fp64}", flush=True)
    t = dev_state(gpu, np.array([1, 2, 3, 4], np.complex128), fp64)
    o, p = M.measure_collapse(t.data_ptr(), 2, 0, 0.5, fp64)
    assert o == 1 and abs(p - 2 / 3) < 1e-14
    o2, p2 = M.measure_collapse(t.data_ptr(), 2, 1, 0.9, fp64)
    assert o2 == 0 and abs(p2 - 0.2) < 1e-14
    psi = host_state(t, fp64)
    assert abs(np.vdot(psi, psi).real - 1) < (1e-14 if fp64 else 1e-6)
    with pytest.raises(M.MiptHipError):
        M.measure_collapse(t.data_ptr(), 2, 2, 0.5, fp64)


@pytest.mark.parametrize("n", [6, 11])
def test_measure_collapse_matches_oracle(M, gpu, n):
    rng = np.random.default_rng(n)
    for trial in range(6):
        psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        psi /= np.linalg.norm(psi)
        q = int(rng.integers(0, n))
        u = float(rng.random())
        t = dev_state(gpu, psi, True)
        o, p = M.measure_collapse(t.data_ptr(), n, q, u, True)
        ref = psi.copy()
        st, oo, pp = O.measure_collapse(ref, n, q, u)
        assert o == oo and abs(p - pp) < 1e-13
        assert np.abs(host_state(t, True) - ref).max() < 1e-13


@pytest.mark.parametrize("n,kept", [(8, [0, 1, 2]), (8, [3, 7, 1]), (10, [0, 1, 2, 3, 4]), (12, [5, 0, 9, 2]),
                                    (14, list(range(7))), (16, list(range(8))), (16, [1]), (13, [12, 0, 6])])
def test_entropy_rdm_renyi_match_oracle(M, gpu, n, kept):
    print(f"entropy n={n} kept={kept}", flush=True)
    psi, _, _, _ = O.simulate(n, 2 * n, 0.1, 1000 + n)
    want = O.entropy(psi, n, kept)
    want2 = O.entropy(psi, n, kept, order=2)
    t32 = dev_state(gpu, psi)
    t64 = dev_state(gpu, psi, True)
    # FP32-MFMA Gram of a complex64 state: within 1e-5 bits of FP64
    assert abs(M.entropy(t32.data_ptr(), n, kept) - want) < 1e-5
    assert abs(M.entropy(t64.data_ptr(), n, kept, fp64=True) - want) < 1e-10
    assert abs(M.entropy(t32.data_ptr(), n, kept, order=2) - want2) < 1e-5
    assert abs(M.entropy(t64.data_ptr(), n, kept, fp64=True, order=2) - want2) < 1e-10
    if len(kept) <= 8:
        rho = O.rdm(psi, n, kept)
        assert np.abs(M.rdm_subsystem(t64.data_ptr(), n, kept, fp64=True) - rho).max() < 1e-12
        assert np.abs(M.rdm_subsystem(t32.data_ptr(), n, kept) - rho).max() < 2e-6


@pytest.mark.parametrize("n,kept", [(12, list(range(6))), (16, list(range(8))), (20, list(range(10)))])
def test_fp32_mfma_gram_within_its_tolerance(M, gpu, n, kept):
    # the FP32-MFMA Gram (selected per process by mipt_hip_set_gram_precision;
    # the default is FP64 MFMA): half-chain entropies within 2e-5 bits of the
    # FP64 oracle and RDM entries within 2e-6; the default restored after
    psi, _, _, _ = O.simulate(n, 2 * n, 0.1, 3000 + n)
    want = O.entropy(psi, n, kept)
    t = dev_state(gpu, psi)
    M.set_gram_precision(32)
    try:
        s32 = M.entropy(t.data_ptr(), n, kept)
        r32 = M.rdm_subsystem(t.data_ptr(), n, kept[:6])
    finally:
        M.set_gram_precision(64)
    s64 = M.entropy(t.data_ptr(), n, kept)
    print(f"  n={n}: FP32 Gram error {abs(s32 - want):.2e}, FP64 Gram error {abs(s64 - want):.2e}", flush=True)
    assert abs(s32 - want) < 2e-5
    assert abs(s64 - want) < 1e-5
    assert np.abs(r32 - O.rdm(psi, n, kept[:6])).max() < 2e-6
    with pytest.raises(M.MiptHipError):
        M.set_gram_precision(16)


def test_entropy_closed_forms_gpu(M, gpu):
    # Bell pairs across the half cut -> exactly 3 bits; product state -> 0
    n = 8
    psi = np.zeros(1 << n, np.complex128)
    for bits in range(8):
        idx = 0
        for (a, b), v in zip([(0, 7), (3, 4), (1, 5)], [(bits >> j) & 1 for j in range(3)]):
            if v:
                idx |= (1 << a) | (1 << b)
        psi[idx] = 1
    psi /= np.linalg.norm(psi)
    t = dev_state(gpu, psi)
    assert abs(M.entropy(t.data_ptr(), n, [0, 1, 2, 3]) - 3.0) < 1e-6
    z = dev_state(gpu, O.zero_state(n))
    assert M.entropy(z.data_ptr(), n, [0, 1, 2, 3]) == 0.0
    assert M.entropy(z.data_ptr(), n, [0, 1, 2, 3], order=2) == 0.0
    ghz = np.zeros(16, np.complex128)
    ghz[0] = ghz[15] = 1 / math.sqrt(2)
    g = dev_state(gpu, ghz)
    i3, terms = M.tmi(g.data_ptr(), 4, 4, 1)
    assert abs(i3 - 1.0) < 1e-6


@pytest.mark.parametrize("n", [12, 16])
def test_tmi_and_rho1_match_oracle(M, gpu, n):
    psi, _, _, _ = O.simulate(n, 2 * n, 0.05, 77 + n)
    t = dev_state(gpu, psi)
    i3, terms = M.tmi(t.data_ptr(), n)
    assert abs(i3 - O.tmi(psi, n)) < 1e-5
    for q in (0, n // 2, n - 1):
        r = M.rho1(t.data_ptr(), n, q)
        assert np.abs(r - O.rho1(psi, n, q)).max() < 1e-6


# ------------------------------------------------------------------ plain engine

@pytest.mark.parametrize("fp64,block", [(False, 2), (False, 4), (True, 4), (True, 2)])
def test_plain_engine_bond_apply(M, fp64, block):
    n = 10
    gates, _, _ = O.build_layers(n, 2, 0.0, 5)
    rng = np.random.default_rng(0)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    e = M.PlainEngine(fp64)
    e.reset(n)
    e.set_from_host(psi)
    bonds = M.layer_bonds(n, 1, True)
    e.apply_bonds(bonds, gates[:len(bonds)], block)
    ref = psi.copy()
    for (q0, q1), g in zip(bonds, gates):
        O.apply_dense(ref, n, [q0, q1], g)
    tol = 1e-12 if fp64 else 2e-6
    assert np.abs(e.to_host() - ref).max() < tol


@pytest.mark.parametrize("fp64", [False, True])
def test_plain_trajectory_replay(M, fp64):
    n, layers, p = 10, 16, 0.3
    print(f"plain replay fp64={fp64}", flush=True)
    psi, want, p1, (g, f, u) = O.simulate(n, layers, p, 42)
    e = M.PlainEngine(fp64)
    got = M.run_plain_trajectory(e, n, g, f, u, True, 4, 0, 10)
    check_outcomes(got, want, p1, u, f, "plain")
    tol = 1e-11 if fp64 else 1e-5
    assert np.abs(e.to_host() - psi).max() < tol


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("n,layers,p,closed,grid_x,blk", [(10, 16, 0.3, True, 0, 4), (12, 12, 0.2, False, 0, 5),
                                                         (12, 8, 0.15, False, 4, 4), (16, 6, 0.1, True, 0, 2)])
def test_plain_trajectory_one_call_matches_loop_and_oracle(M, fp64, n, layers, p, closed, grid_x, blk):
    # mipt_hip_engine_run_trajectory: the whole try_cusv loop in one C-ABI call,
    # against the layer-by-layer host loop (same outcomes, bit-identical state)
    # and against the oracle
    print(f"plain one-call n={n} grid_x={grid_x} fp64={fp64}", flush=True)
    if grid_x:
        psi, want, p1, f, u = O.simulate_grid(grid_x, n // grid_x, layers, p, 77, closed)
        g = []
        for t in range(layers):
            for b, _ in enumerate(O.grid_layer_bonds(grid_x, n // grid_x, t, closed)):
                gg = np.zeros(32)
                O.lib().orc_philox_haar(77, t, b, gg.ctypes.data_as(O._dp))
                g.append((gg[0::2] + 1j * gg[1::2]).reshape(4, 4))
        g = np.array(g)
    else:
        psi, want, p1, (g, f, u) = O.simulate(n, layers, p, 77, closed)
    e1, e2 = M.PlainEngine(fp64), M.PlainEngine(fp64)
    got1 = M.run_plain_trajectory_native(e1, n, g, f, u, closed, blk, 0, 10, grid_x)
    got2 = M.run_plain_trajectory(e2, n, g, f, u, closed, blk, 0, 10, grid_x)
    assert np.array_equal(got1, got2)
    assert np.array_equal(e1.to_host(), e2.to_host())
    check_outcomes(got1, want, p1, u, f, "plain one-call")
    tol = 1e-11 if fp64 else 1e-5
    if not (np.abs(u - p1) < FLAG_EPS).any():
        assert np.abs(e1.to_host() - psi).max() < tol
    e1.close()
    e2.close()


def test_plain_trajectory_one_call_joint_mode(M):
    n, layers = 8, 10
    gates, flags, _ = O.build_layers(n, layers, 0.4, 9)
    rng = np.random.default_rng(1)
    unif = rng.random((layers, n))
    ref = O.zero_state(n)
    want, _ = O.run_layers(ref, n, gates, flags, unif, mode=1, chunk_bits=3)
    e = M.PlainEngine(True)
    got = M.run_plain_trajectory_native(e, n, gates, flags, unif, True, 4, 1, 3)
    assert np.array_equal(got, want)
    assert np.abs(e.to_host() - ref).max() < 1e-11


def test_plain_joint_mode_matches_oracle(M):
    # engine.hpp:293-368 joint chunk sampling (one uniform per chunk)
    n, layers = 8, 10
    gates, flags, _ = O.build_layers(n, layers, 0.4, 9)
    rng = np.random.default_rng(1)
    unif = rng.random((layers, n))
    ref = O.zero_state(n)
    want, _ = O.run_layers(ref, n, gates, flags, unif, mode=1, chunk_bits=3)
    e = M.PlainEngine(True)
    got = M.run_plain_trajectory(e, n, gates, flags, unif, True, 4, 1, 3)
    assert np.array_equal(got, want)
    assert np.abs(e.to_host() - ref).max() < 1e-11


# ------------------------------------------------------------------ batched engine

def replay_batch(M, cases, tile_bits=0, threads=0, arith=0):
    """cases: list of (n, layers, p, seed, closed); returns (batch, oracle results)."""
    specs = [M.Spec(n, layers, p, seed, closed) for (n, layers, p, seed, closed) in cases]
    b = M.Batch(specs, tile_bits=tile_bits, threads=threads, bond_arith=arith)
    refs = []
    for i, (n, layers, p, seed, closed) in enumerate(cases):
        psi, out, p1, (g, f, u) = O.simulate(n, layers, p, seed, closed)
        b.set_layers(i, g, f, u)
        refs.append((psi, out, p1, f, u))
    b.evolve()
    b.synchronize()
    return b, refs


@pytest.mark.parametrize("tile_bits,threads,arith", [(0, 0, 0), (14, 0, 0), (13, 0, 1), (13, 0, 2), (13, 0, 3),
                                                     (13, 128, 0), (12, 0, 0)])
def test_batch_replay_mixed_lengths(M, tile_bits, threads, arith):
    cases = [(4, 8, 0.3, 1, True), (8, 12, 0.2, 2, True), (12, 14, 0.16, 3, False), (14, 10, 0.3, 4, True),
             (16, 8, 0.2, 5, True), (16, 9, 0.5, 6, False), (18, 6, 0.3, 7, True), (20, 5, 0.25, 8, True)]
    print(f"batch replay tile_bits={tile_bits} threads={threads} bond_arith={arith}", flush=True)
    b, refs = replay_batch(M, cases, tile_bits, threads, arith)
    for i, (psi, out, p1, f, u) in enumerate(refs):
        got, flagged, gp1, _ = b.outcomes(i)
        check_outcomes(got, out, p1, u, f, f"traj {i} L={cases[i][0]}")
        m = f == 1
        assert np.abs(gp1[m] - p1[m]).max(initial=0) < 1e-5
        assert np.all(flagged[m] == (np.abs(u[m] - gp1[m]) < FLAG_EPS))
        if not np.any(flagged):
            assert np.abs(b.get_state(i) - psi).max() < 1e-5, f"traj {i}"


def test_batch_overflow_chunks_and_edge_rates(M):
    # p=1 measures every site (16 > one 10-site histogram chunk); p=0 none.
    cases = [(16, 6, 1.0, 11, True), (16, 6, 0.0, 12, True), (10, 6, 1.0, 13, True), (20, 4, 0.6, 14, True)]
    b, refs = replay_batch(M, cases)
    for i, (psi, out, p1, f, u) in enumerate(refs):
        got, flagged, _, _ = b.outcomes(i)
        check_outcomes(got, out, p1, u, f, f"traj {i}")
        if not np.any(flagged):
            assert np.abs(b.get_state(i) - psi).max() < 1e-5


def test_batch_device_stream_matches_oracle_stream(M):
    # gates, flags and uniforms drawn on device (Philox) equal the oracle's
    cases = [(10, 10, 0.2, 101), (16, 8, 0.25, 102), (18, 6, 0.16, 103)]
    specs = [M.Spec(n, layers, p, seed) for (n, layers, p, seed) in cases]
    b = M.Batch(specs)
    b.generate()
    b.evolve()
    b.synchronize()
    for i, (n, layers, p, seed) in enumerate(cases):
        psi, out, p1, (g, f, u) = O.simulate(n, layers, p, seed)
        got, flagged, _, _ = b.outcomes(i)
        check_outcomes(got, out, p1, u, f, f"gen {i}")
        if not np.any(flagged):
            assert np.abs(b.get_state(i) - psi).max() < 1e-5


def test_batch_observables_match_oracle(M):
    cases = [(12, 12, 0.1, 21, True), (16, 16, 0.1, 22, True), (20, 6, 0.1, 23, True)]
    b, refs = replay_batch(M, cases)
    for i, (psi, out, p1, f, u) in enumerate(refs):
        n = cases[i][0]
        ob = b.observables(i)
        assert abs(ob["s_half"] - O.half_chain_entropy(psi, n)) < 1e-5
        assert abs(ob["tmi"] - O.tmi(psi, n)) < 1e-5
        assert math.isnan(ob["probe_entropy"])


def test_batch_probe_bell_pair(M):
    # probe qubit at index L Bell-paired with site L/2 (probed.hpp:25-33): with
    # p = 0 the probe stays maximally entangled: S = ln 2, Bloch = 0.
    specs = [M.Spec(8, 8, 0.0, 31, True, probe_site=4), M.Spec(16, 4, 0.0, 32, True, probe_site=8),
             M.Spec(10, 10, 0.3, 33, True, probe_site=5)]
    b = M.Batch(specs)
    b.generate()
    b.evolve()
    b.synchronize()
    for i in range(2):
        ob = b.observables(i, want_tmi=False)
        assert abs(ob["probe_entropy"] - math.log(2)) < 1e-6
        assert np.abs(ob["bloch"]).max() < 1e-6
    # measured case against the oracle (probe qubit = index L, never gated)
    n = 10
    g, f, u = O.build_layers(n, 10, 0.3, 33)
    psi = np.zeros(1 << (n + 1), np.complex128)
    psi[0] = psi[(1 << 5) | (1 << n)] = 1 / math.sqrt(2)
    # evolve the (n+1)-qubit state: gates act on 0..n-1 only
    outs = np.full((10, n), -1, np.int32)
    gi = 0
    for t in range(10):
        for (q0, q1) in O.layer_bonds(n, t, True):
            O.apply_dense(psi, n + 1, [q0, q1], g[gi])
            gi += 1
        for s in range(n):
            if f[t, s]:
                st, o, _ = O.measure_collapse(psi, n + 1, s, u[t, s])
                outs[t, s] = o
    got, flagged, _, _ = b.outcomes(2)
    if not flagged.any():
        assert np.array_equal(got, outs)
        r = O.rho1(psi, n + 1, n)
        ob = b.observables(2, want_tmi=False)
        assert abs(ob["probe_entropy"] - O.rho1_entropy_nats(r)) < 1e-5
        assert np.abs(ob["bloch"] - O.bloch_from_rho1(r)).max() < 1e-5


def test_batch_probe_attach_after_scrambling(M):
    # the decoder workload's protocol: one shared gate realisation (donor
    # trajectory), unmeasured scrambling layers, then reset(site) + Bell pair
    # with the probe (probed/kernels.hpp:68-77), then measured layers.
    # L=8 runs in the small kernel, L=16 (+ probe = 17 qubits) in tile sweeps.
    specs = [M.Spec(8, 12, 0.2, 7, True, probe_site=4, p_zero_layers=4),
             M.Spec(8, 12, 0.3, 1001, True, probe_site=4, p_zero_layers=4, shared_gates=0),
             M.Spec(16, 10, 0.2, 7, True, probe_site=8, p_zero_layers=4),
             M.Spec(16, 10, 0.25, 1002, True, probe_site=8, p_zero_layers=4, shared_gates=2)]
    donors = [None, 7, None, 7]
    print("batch probe attach after scrambling", flush=True)
    b = M.Batch(specs)
    b.generate()
    b.evolve()
    b.synchronize()
    bloch = b.probe_bloch()
    for i, s in enumerate(specs):
        psi, outs, p1, att, (g, f, u) = O.probe_trajectory(s.L, s.layers, s.p, s.seed, s.probe_site,
                                                           s.p_zero_layers, donors[i])
        got, flagged, _, _ = b.outcomes(i)
        check_outcomes(got, outs, p1, u, f, f"probe traj {i}")
        if not flagged.any():
            assert np.abs(b.get_state(i) - psi).max() < 1e-5, f"traj {i}"
            want = O.bloch_from_rho1(O.rho1(psi, s.L + 1, s.L))
            assert np.abs(bloch[i] - want).max() < 1e-5
    with pytest.raises(M.MiptHipError):
        b.records(4, 10)  # mixed chain lengths
    b2 = M.Batch([M.Spec(16, 10, 0.2, 7, True, probe_site=8, p_zero_layers=4),
                  M.Spec(16, 10, 0.25, 1002, True, probe_site=8, p_zero_layers=4, shared_gates=0)])
    b2.generate()
    b2.evolve()
    b2.synchronize()
    rec = b2.records(4, 10)
    for i in range(2):
        got, _, _, _ = b2.outcomes(i)
        assert np.array_equal(rec[i], got[4:10].astype(np.int8))
        assert np.array_equal(got, b.outcomes(i + 2)[0])


def test_batch_grid_matches_oracle(M):
    # 2D grids with the four-colour layering: 4 rows x 5 columns (20 qubits,
    # tile sweeps with windows chosen to cover the vertical bonds), 3 x 4
    # (small kernel) and 5 x 4 with closed rows; at p = 0.03 most layers of
    # the 4 x 5 grid fuse into the next (its compiled fused sweeps)
    cases = [(5, 4, 8, 0.12, 2001, False), (4, 3, 8, 0.2, 2002, False), (4, 5, 6, 0.3, 2003, True),
             (5, 4, 12, 0.03, 2004, False)]
    specs = [M.Spec(gx * gy, layers, p, seed, closed, grid_x=gx) for (gx, gy, layers, p, seed, closed) in cases]
    print("batch 2D grids", flush=True)
    b = M.Batch(specs)
    b.generate()
    b.evolve()
    b.synchronize()
    for i, (gx, gy, layers, p, seed, closed) in enumerate(cases):
        n = gx * gy
        psi, outs, p1, f, u = O.simulate_grid(gx, gy, layers, p, seed, closed)
        got, flagged, _, _ = b.outcomes(i)
        check_outcomes(got, outs, p1, u, f, f"grid {gx}x{gy}")
        if not flagged.any():
            assert np.abs(b.get_state(i) - psi).max() < 1e-5, f"grid {i}"
            ob = b.observables(i)
            assert abs(ob["s_half"] - O.half_chain_entropy(psi, n)) < 1e-5
            assert abs(ob["tmi"] - O.tmi(psi, n)) < 1e-5
        # the exported draws replay on the plain engine too
        g, fl, un = b.export_layers(i)
        eng = M.PlainEngine(False)
        want = M.run_plain_trajectory(eng, n, g, fl, un, closed, 4, 0, 10, grid_x=gx)
        assert np.array_equal(want[fl == 1], outs[fl == 1]) or flagged.any()
        eng.close()


def test_batch_observables_all_matches_single_calls_and_oracle(M):
    # batched eigensolves give the same S(L/2) and per-cycle TMI as the
    # one-at-a-time entry points, and the oracle's within 1e-5
    cases = [(12, 10, 0.1, 41), (12, 8, 0.2, 42), (16, 8, 0.15, 43), (20, 4, 0.1, 44), (16, 6, 0.0, 45),
             (8, 8, 0.2, 46), (8, 6, 0.1, 47), (10, 6, 0.2, 48)]
    specs = [M.Spec(n, layers, p, seed) for (n, layers, p, seed) in cases]
    print("batch observables_all", flush=True)
    b = M.Batch(specs)
    b.generate()
    b.evolve()
    b.synchronize()
    sh, tm = b.observables_all(0, 3)
    for i, (n, layers, p, seed) in enumerate(cases):
        ob = b.observables(i)
        assert abs(sh[i] - ob["s_half"]) < 1e-9
        for c in range(3):
            assert abs(tm[i, c] - b.tmi(i, c)) < 1e-9
        psi, *_ = O.simulate(n, layers, p, seed)
        got, flagged, _, _ = b.outcomes(i)
        if not flagged.any():
            assert abs(sh[i] - O.half_chain_entropy(psi, n)) < 1e-5
            for c in range(3):
                assert abs(tm[i, c] - O.tmi(psi, n, n // 4, c)) < 1e-5


def test_batch_autotune_then_evolve_matches_oracle(M):
    cases = [(16, 6, 0.2, 81), (18, 6, 0.25, 82), (20, 5, 0.15, 83), (10, 6, 0.3, 84)]
    specs = [M.Spec(n, layers, p, seed) for (n, layers, p, seed) in cases]
    print("batch autotune", flush=True)
    b = M.Batch(specs)
    b.generate()
    best, plan, ms = b.autotune(3, with_plan=True)
    print("autotune choice", best, "plan", plan, "ms", ms, flush=True)
    assert 0 <= best < 6 and ms[best] > 0 and 0 <= plan < 4 and ms[6:].min() > 0
    b.evolve()
    b.synchronize()
    for i, (n, layers, p, seed) in enumerate(cases):
        psi, out, p1, (g, f, u) = O.simulate(n, layers, p, seed)
        got, flagged, _, _ = b.outcomes(i)
        check_outcomes(got, out, p1, u, f, f"tuned {i}")
        if not flagged.any():
            assert np.abs(b.get_state(i) - psi).max() < 1e-5


def test_batch_graph_replays_match_eager(M):
    # graph replays (with eager work and synchronisation in between) give the
    # eager run's outcomes and states exactly
    cases = [(16, 6, 0.2, 91), (20, 4, 0.3, 92), (10, 6, 0.25, 93), (18, 5, 0.5, 94)]
    specs = [M.Spec(n, layers, p, seed) for (n, layers, p, seed) in cases]
    print("batch graph replays", flush=True)
    b = M.Batch(specs)
    b.set_graph(True)
    ref = None
    for rep in range(4):  # eager, capture + launch, replay, replay
        b.reset()
        b.generate()
        b.evolve()
        b.synchronize()
        got = [(b.outcomes(i)[0], b.get_state(i)) for i in range(len(cases))]
        b.observables(0)  # eager work between replays
        if ref is None:
            ref = got
            continue
        for i in range(len(cases)):
            assert np.array_equal(got[i][0], ref[i][0]), (rep, i)
            assert np.array_equal(got[i][1], ref[i][1]), (rep, i)
