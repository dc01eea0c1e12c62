"""The large single-size workload: L=26 chains (2^26 complex64 amplitudes,
512 MiB per state), three sweeps per layer, and the 8192 x 8192 half-chain
RDM.

Full-size parity (the FP64 oracle is too slow at 2^26 amplitudes for a
test): the batched engine against the plain engine on the same device draws
(the plain engine is itself pinned to the oracle at small L), outcomes
bit-exact except flagged draws, amplitudes within 1e-5; the 8192^2 and 4096^2
entropy paths against closed forms of Bell-pair states (S(X) = number of
pairs with exactly one end in X), within 1e-5 bits.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 26


@pytest.fixture(scope="module")
def M():
    import mipt_amd
    mipt_amd.lib()
    return mipt_amd


def test_l26_batch_matches_plain_engine(M):
    print("L=26 batch vs plain, 2 trajectories x 6 layers", flush=True)
    specs = [M.Spec(N, 6, 0.16, 2601), M.Spec(N, 6, 0.3, 2602)]
    b = M.Batch(specs)
    b.generate()
    b.evolve()
    b.synchronize()
    eng = M.PlainEngine(False)
    try:
        for i, s in enumerate(specs):
            g, f, u = b.export_layers(i)
            got, flagged, _, _ = b.outcomes(i)
            want = M.run_plain_trajectory(eng, N, g, f, u, True, 4, 0, 10)
            bad = (got != want) & (flagged == 0)
            assert not bad.any(), f"traj {i}: outcome mismatch at {np.argwhere(bad)[:4]}"
            assert np.array_equal(got >= 0, f == 1)
            if not flagged.any():
                psi_b = b.get_state(i)
                psi_p = eng.to_host()
                assert np.abs(psi_b - psi_p).max() < 1e-5
                assert abs(eng.squared_norm() - 1.0) < 1e-5
    finally:
        eng.close()


def bell_pairs_state(n, pairs):
    """Product of Bell pairs (a, b) (|00> + |11>)/sqrt(2); other qubits |0>."""
    idx = np.zeros(1, np.int64)
    for a, bq in pairs:
        idx = np.concatenate([idx, idx | (1 << a) | (1 << bq)])
    psi = np.zeros(1 << n, np.complex64)
    psi[idx] = 1.0 / math.sqrt(len(idx))
    return psi


def cut_count(pairs, region):
    r = set(region)
    return sum((a in r) != (b in r) for a, b in pairs)


def test_l26_half_chain_rdm_8192_closed_form(M):
    # 13 pairs (i, i+13): the half chain is maximally entangled, S(13) = 13
    pairs = [(i, i + 13) for i in range(13)]
    print("L=26 half-chain 8192^2 RDM", flush=True)
    psi = bell_pairs_state(N, pairs)
    eng = M.PlainEngine(False)
    try:
        eng.reset(N)
        eng.set_from_host(psi)
        s = M.entropy(eng.data, N, list(range(13)))
        assert abs(s - 13.0) < 1e-5
        # TMI with 6-site blocks (A=0..5, B=6..11, C=12..17, D=18..25)
        blocks = [list(range(0, 6)), list(range(6, 12)), list(range(12, 18)), list(range(18, 26))]
        A, B, C, D = blocks
        want = (cut_count(pairs, A) + cut_count(pairs, B) + cut_count(pairs, C) + cut_count(pairs, D)
                - cut_count(pairs, A + B) - cut_count(pairs, A + C) - cut_count(pairs, B + C))
        got, terms = M.tmi(eng.data, N, N, 6, 0)
        assert abs(got - want) < 1e-5
    finally:
        eng.close()
