"""Several reference qubits (the probe protocols' two-, three- and
four-reference samples): the oracle's joint RDM and information measures
pinned by the existing RDM routine and by closed forms; the library's host
information routine (mipt_hip_probe_information) against the oracle's.

Closed forms (nats): maximally mixed references give S = k ln 2 and no
information; a Bell pair gives S(AB) = 0, I(A:B) = 2 ln 2, negativity 1/2,
log-negativity ln 2; a 4-qubit GHZ state gives I2 = ln 2, I3 = ln 2,
I4 = 2 ln 2; its 3-qubit marginal has I3 = ln 2 and no negativity; a
3-qubit GHZ state has I3 = 0, mean pair MI ln 2, cut negativity 1/2,
purities 1 and 1/2; two Bell pairs (AB)(CD) give mean I2 = (2/3) ln 2.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

LN2 = math.log(2.0)


@pytest.fixture(scope="module")
def M():
    import mipt_amd
    mipt_amd.lib()
    return mipt_amd


def ket(k, amps):
    v = np.zeros(1 << k, np.complex128)
    for i, a in amps.items():
        v[i] = a
    return np.outer(v, v.conj())


GHZ3 = ket(3, {0: 2 ** -0.5, 7: 2 ** -0.5})
GHZ4 = ket(4, {0: 2 ** -0.5, 15: 2 ** -0.5})
BELL = ket(2, {0: 2 ** -0.5, 3: 2 ** -0.5})
CASES = [
    (np.eye(2) / 2, dict(sq=LN2)),
    (np.eye(4) / 4, dict(sq=2 * LN2, i2=0, negativity=0, log_negativity=0)),
    (BELL, dict(sq=0, i2=2 * LN2, negativity=0.5, log_negativity=LN2)),
    (GHZ3, dict(tmi=0, average_mi=LN2, min_bipartite_negativity=0.5, joint_purity=1, mean_single_purity=0.5)),
    (O.partial_trace_bits(GHZ4, 4, [0, 1, 2]), dict(tmi=LN2, average_mi=LN2, min_bipartite_negativity=0,
                                                   joint_purity=0.5, mean_single_purity=0.5)),
    (np.eye(16) / 16, dict(sq=4 * LN2, i2=0, i3=0, i4=0)),
    (GHZ4, dict(sq=0, i2=LN2, i3=LN2, i4=2 * LN2)),
    (np.kron(BELL, BELL), dict(sq=0, i2=2 / 3 * LN2, i3=0, i4=0)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_information_closed_forms(M, case):
    rho, want = CASES[case]
    for got in (O.probe_information(rho), M.probe_information(rho)):
        for key, v in want.items():
            assert abs(got[key] - v) < 1e-12, (case, key, got[key], v)


def test_reference_rdm_and_partial_trace_match_rdm_routine():
    # reduced-basis bit j = reference j (qubit n + j): the same convention as
    # the entropy path's RDM of a kept list
    rng = np.random.default_rng(5)
    n = 5
    for k in (1, 2, 3, 4):
        psi = rng.normal(size=1 << (n + k)) + 1j * rng.normal(size=1 << (n + k))
        psi /= np.linalg.norm(psi)
        rho = O.reference_rdm(psi, n, k)
        assert np.abs(rho - O.rdm(psi, n + k, list(range(n, n + k)))).max() < 1e-13
        for keep in ([0], [k - 1], [0, k - 1], list(range(k))):
            keep = sorted(set(keep))
            want = O.rdm(psi, n + k, [n + j for j in keep])
            assert np.abs(O.partial_trace_bits(rho, k, keep) - want).max() < 1e-13


def test_cut_negativity_is_the_partial_transpose_trace_norm():
    # brute force: transpose the chosen party's bit explicitly
    rng = np.random.default_rng(6)
    for k, j in ((2, 0), (2, 1), (3, 1), (3, 2)):
        D = 1 << k
        a = rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))
        rho = a @ a.conj().T
        rho /= np.trace(rho).real
        m = 1 << j
        pt = np.array([[rho[(r & ~m) | (c & m), (c & ~m) | (r & m)] for c in range(D)] for r in range(D)])
        want = 0.5 * (np.abs(np.linalg.eigvalsh(pt)).sum() - 1)
        assert abs(O.cut_negativity(rho, k, j) - max(want, 0.0)) < 1e-12


def test_library_information_matches_oracle(M):
    rng = np.random.default_rng(7)
    for k in (1, 2, 3, 4):
        for rank in (1, 2, 1 << k):
            D = 1 << k
            a = rng.normal(size=(D, rank)) + 1j * rng.normal(size=(D, rank))
            rho = a @ a.conj().T
            rho /= np.trace(rho).real
            want, got = O.probe_information(rho), M.probe_information(rho)
            for key in want:
                assert abs(got[key] - want[key]) < 1e-10, (k, rank, key)


def test_library_information_rejects_bad_input(M):
    with pytest.raises(M.MiptHipError):
        M.probe_information(np.zeros((4, 4)))
    with pytest.raises(M.MiptHipError):
        M.probe_information(np.eye(32) / 32)


def bell_swap_gate():
    """H on bit 0 after CNOT(bit 0 -> bit 1): maps the Bell basis to the
    computational one, so measuring both sites is a Bell measurement."""
    h = np.array([[1, 1], [1, -1]]) / math.sqrt(2)
    cnot = np.zeros((4, 4))
    for b0 in (0, 1):
        for b1 in (0, 1):
            cnot[b0 | ((b1 ^ b0) << 1), b0 | (b1 << 1)] = 1
    return np.kron(np.eye(2), h) @ cnot


def swap_layers(n, s0, layers=2):
    """Identity gates except the bond (s0, s0+1) of layer 1, whose two sites
    are measured in that layer (uniforms 0.3, 0.7)."""
    gates = []
    for t in range(layers):
        for (q0, q1) in O.layer_bonds(n, t, True):
            gates.append(bell_swap_gate() if (t == 1 and (q0, q1) == (s0, s0 + 1)) else np.eye(4))
    flags = np.zeros((layers, n), np.uint8)
    unif = np.zeros((layers, n))
    flags[1, s0] = flags[1, s0 + 1] = 1
    unif[1, s0], unif[1, s0 + 1] = 0.3, 0.7
    return np.array(gates, np.complex128), flags, unif


def test_entanglement_swapping_oracle():
    # references Bell-paired with sites s0, s0+1 at t = 0; a Bell measurement
    # of the two sites leaves the references in a Bell state
    for n, s0 in ((4, 1), (8, 3)):
        psi, *_ = O.references_trajectory(n, 0, 0.0, 1, [s0, s0 + 1], 0)
        assert np.abs(O.reference_rdm(psi, n, 2) - np.eye(4) / 4).max() < 1e-15
        g, f, u = swap_layers(n, s0)
        gi = 0
        for t in range(2):
            for (q0, q1) in O.layer_bonds(n, t, True):
                O.apply_dense(psi, n + 2, [q0, q1], g[gi])
                gi += 1
            for s in range(n):
                if f[t, s]:
                    O.measure_collapse(psi, n + 2, s, u[t, s])
        info = O.probe_information(O.reference_rdm(psi, n, 2))
        assert abs(info["sq"]) < 1e-12 and abs(info["i2"] - 2 * LN2) < 1e-12
        assert abs(info["negativity"] - 0.5) < 1e-12


def test_attach_after_scrambling_oracle_invariants():
    # after the attach every reference is maximally mixed (the reset makes the
    # site |0>, h + cx make the pair maximally entangled) and the state is normalised
    for k in (2, 3, 4):
        sites = [1, 3, 5, 7][:k]
        psi, outs, p1, att, _ = O.references_trajectory(8, 5, 0.0, 11, sites, 5)
        assert len(att) == 0  # the attach layer is layer 5 = the end
        psi, outs, p1, att, _ = O.references_trajectory(8, 6, 0.0, 11, sites, 5)
        assert len(att) == k and abs(np.vdot(psi, psi).real - 1) < 1e-12
        rho = O.reference_rdm(psi, 8, k)
        for j in range(k):
            assert np.abs(O.partial_trace_bits(rho, k, [j]) - np.eye(2) / 2).max() < 1e-12
