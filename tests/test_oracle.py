"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each check is chosen so that a plausible mistake in the oracle fails it:
known-answer tests from the reference's own test files, the reference's host
code compiled into oracle/_ref (when the reference tree is present), closed
forms (Bell pairs, GHZ, Page values), brute force on tiny inputs, and the
published Philox known-answer vectors.  Paths are relative to the reference
repository root.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O

class _LazyRef:
    """oracle/_ref/libref.so, loaded on first use (GPU test runs never load it)."""

    def __getattr__(self, name):
        return getattr(O.ref_lib(), name)


REF = _LazyRef() if O.ref_available() else None
needs_ref = pytest.mark.skipif(REF is None, reason="reference tree absent (oracle/_ref not built)")


# ------------------------------------------------------------------ seeding

def test_splitmix_and_trajectory_seed_against_reference_code():
    # types.hpp:262-273; compiled reference when available, else the
    # published splitmix64 first output for state 0 (0xe220a8397b1dcdaf).
    assert O.lib().orc_splitmix64(0) == 0xE220A8397B1DCDAF
    if REF is not None:
        for v in [0, 1, 12345, 2**63 + 7, 2**64 - 1]:
            assert O.lib().orc_splitmix64(v) == REF.ref_splitmix64(v)
        for m, i in [(0, 0), (1, 5), (2024, 4095), (2**40, 17)]:
            assert O.lib().orc_trajectory_seed(m, i) == REF.ref_trajectory_seed(m, i)


@needs_ref
def test_linspace_against_reference_code():
    for a, b, c in [(0.0, 0.3, 16), (0.1, 0.22, 3), (0.05, 0.05, 1), (0.0, 1.0, 7)]:
        x = np.zeros(c)
        y = np.zeros(c)
        O.lib().orc_linspace(a, b, c, O._d(x))
        REF.ref_linspace(a, b, c, O._d(y))
        assert np.array_equal(x, y)


def test_philox_known_answer_vectors():
    # Random123 kat_vectors for philox4x32_10.
    assert O.philox4x32_10([0, 0], [0, 0, 0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox4x32_10([0xFFFFFFFF, 0xFFFFFFFF], [0xFFFFFFFF] * 4) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox4x32_10([0xA4093822, 0x299F31D0], [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


# ------------------------------------------------------------------ layers and Haar gates

def test_brickwork_bond_layout():
    # haar.hpp:226-243: even layers (0,1),(2,3)..; odd closed layers end with (n-1,0).
    assert O.layer_bonds(8, 0, True) == [(0, 1), (2, 3), (4, 5), (6, 7)]
    assert O.layer_bonds(8, 1, True) == [(1, 2), (3, 4), (5, 6), (7, 0)]
    assert O.layer_bonds(8, 1, False) == [(1, 2), (3, 4), (5, 6)]
    assert O.total_bonds(20, 40, True) == 400  # 40 layers x 10 bonds
    assert O.total_bonds(12, 24, False) == 12 * 6 + 12 * 5  # open chain


@pytest.mark.parametrize("stream", ["philox", "mt"])
def test_haar_gates_unitary_and_haar_moments(stream):
    # Unitarity < 1e-12 over 400 bonds (tests/cusv_gate_tests.cpp:128-156);
    # Weingarten moments of U(4): E|U_ij|^2 = 1/4, E|U_ij|^4 = 2/(d(d+1)) = 1/10.
    g, _, _ = O.build_layers(20, 40, 0.0, 7, True, stream)
    eye = np.eye(4)
    err = max(np.abs(u.conj().T @ u - eye).max() for u in g)
    assert err < 1e-12
    allg = []
    for s in range(1, 26):
        gs, _, _ = O.build_layers(20, 40, 0.0, s, True, stream)
        allg.append(gs)
    a = np.abs(np.concatenate(allg)) ** 2  # 10000 gates x 16 entries
    assert abs(a.mean() - 0.25) < 2e-3
    assert abs((a ** 2).mean() - 0.1) < 2e-3
    # no entry is systematically real: phases uniform
    ph = np.angle(np.concatenate(allg)).ravel()
    assert abs(np.mean(np.cos(ph))) < 0.02 and abs(np.mean(np.sin(ph))) < 0.02


def test_flags_are_bernoulli_and_edge_cases():
    _, f, u = O.build_layers(20, 400, 0.16, 3, True, "philox")
    assert abs(f.mean() - 0.16) < 0.01
    assert np.all((u > 0) == (f == 1)) or np.all(u[f == 0] == 0)
    assert np.all((u >= 0) & (u < 1))
    _, f0, _ = O.build_layers(8, 10, 0.0, 3)
    _, f1, _ = O.build_layers(8, 10, 1.0, 3)
    assert f0.sum() == 0 and f1.sum() == 80


def test_streams_are_seed_deterministic_and_distinct():
    a = O.build_layers(12, 6, 0.3, 11)
    b = O.build_layers(12, 6, 0.3, 11)
    c = O.build_layers(12, 6, 0.3, 12)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])


# ------------------------------------------------------------------ gate application

def kron_full(n, targets, m):
    """Dense 2^n x 2^n operator of m on targets (bit j of m's index = targets[j])."""
    dim = 1 << n
    full = np.zeros((dim, dim), np.complex128)
    k = len(targets)
    for col in range(dim):
        a = sum(((col >> t) & 1) << j for j, t in enumerate(targets))
        base = col
        for t in targets:
            base &= ~(1 << t)
        for r in range(1 << k):
            row = base
            for j, t in enumerate(targets):
                if (r >> j) & 1:
                    row |= 1 << t
            full[row, col] += m[r, a]
    return full


@pytest.mark.parametrize("n,targets", [(4, [0, 1]), (5, [3, 1]), (6, [5, 0]), (8, [2, 7, 4]), (7, [6, 1, 0, 3])])
def test_apply_dense_matches_brute_force(n, targets):
    rng = np.random.default_rng(5)
    k = len(targets)
    m = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    want = kron_full(n, targets, m) @ psi
    got = psi.copy()
    O.apply_dense(got, n, targets, m)
    assert np.abs(got - want).max() < 1e-12


@needs_ref
def test_tensor_product_blocking_matches_reference_code():
    # cusv/matrix.hpp:214-235: applying kron(high, low) on low+high targets
    # equals applying the bonds one by one (the blocking invariant).
    rng = np.random.default_rng(1)
    lo = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    hi = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    out = np.zeros((16, 16), np.complex128)
    REF.ref_tensor_product(lo.ctypes.data_as(O._dp), 2, hi.ctypes.data_as(O._dp), 2, out.ctypes.data_as(O._dp))
    assert np.abs(out - np.kron(hi, lo)).max() < 1e-14
    psi = rng.normal(size=1 << 6) + 1j * rng.normal(size=1 << 6)
    a = psi.copy()
    O.apply_dense(a, 6, [0, 1], lo)
    O.apply_dense(a, 6, [4, 5], hi)
    b = psi.copy()
    O.apply_dense(b, 6, [0, 1, 4, 5], out)
    assert np.abs(a - b).max() < 1e-12


# ------------------------------------------------------------------ measurement

def test_sequential_measurement_kat():
    # tests/core_tests.cpp:120-157: psi ~ [1,2,3,4]; q0 with u=0.5 -> 1, P=2/3;
    # q1 with u=0.9 -> 0, P=0.2; joint 4/30; norm 1 after each collapse.
    psi = np.array([1, 2, 3, 4], np.complex128)
    st, o, p = O.measure_collapse(psi, 2, 0, 0.5)
    assert st == 0 and o == 1 and abs(p - 2 / 3) < 1e-14
    assert abs(np.vdot(psi, psi).real - 1) < 1e-14
    st, o2, p2 = O.measure_collapse(psi, 2, 1, 0.9)
    assert st == 0 and o2 == 0 and abs(p2 - 0.2) < 1e-14
    assert abs(p * p2 - 4 / 30) < 1e-14
    assert abs(np.vdot(psi, psi).real - 1) < 1e-14


def test_measurement_argument_errors():
    psi = np.array([1, 0, 0, 0], np.complex128)
    assert O.measure_collapse(psi, 2, 2, 0.5)[0] == -1
    assert O.measure_collapse(psi, 2, 0, 1.0)[0] == -1
    z = np.zeros(4, np.complex128)
    assert O.measure_collapse(z, 2, 0, 0.5)[0] == -2


@needs_ref
def test_sequential_measurement_matches_reference_code():
    rng = np.random.default_rng(2)
    for trial in range(30):
        n = int(rng.integers(1, 9))
        psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        psi /= np.linalg.norm(psi) * (1 + 1e-3 * rng.normal())
        q = int(rng.integers(0, n))
        u = float(rng.random())
        a = psi.copy()
        b = psi.copy()
        st, o, p = O.measure_collapse(a, n, q, u)
        oo, pp, ss = O.ctypes.c_int(), O.ctypes.c_double(), O.ctypes.c_double()
        rs = REF.ref_measure_collapse_host(b.ctypes.data_as(O._dp), n, q, u, O.ctypes.byref(oo), O.ctypes.byref(pp),
                                           O.ctypes.byref(ss))
        assert st == 0 and rs == 0
        assert o == oo.value and p == pp.value
        assert np.array_equal(a, b)


def test_joint_chunked_sampling_equals_sequential_chain_rule():
    # Joint sampling of a layer equals sequential sampling by the chain rule
    # (tests/cusv_equiv.cpp:202-277): same post-measurement state for the
    # same outcomes, and the joint probability is the product of conditionals.
    rng = np.random.default_rng(3)
    n = 6
    for trial in range(20):
        psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        psi /= np.linalg.norm(psi)
        sites = sorted(rng.choice(n, size=int(rng.integers(1, n + 1)), replace=False).tolist())
        a = psi.copy()
        outs = np.zeros(len(sites), np.int32)
        st = O.ctypes.c_int()
        u = np.array([rng.random()])
        s_arr = np.array(sites, np.int32)
        surpr = O.lib().orc_measure_layer_joint(a.ctypes.data_as(O._dp), n, O._i(s_arr), len(sites), 16, O._d(u),
                                                 O._i(outs), O.ctypes.byref(st))
        assert st.value == 0
        # replay the outcomes sequentially with a projector
        b = psi.copy()
        logp = 0.0
        for s, o in zip(sites, outs):
            idx = np.arange(1 << n)
            w1 = np.sum(np.abs(b[(idx >> s) & 1 == 1]) ** 2)
            w = np.sum(np.abs(b) ** 2)
            p = w1 / w if o == 1 else 1 - w1 / w
            logp += math.log(p)
            b[((idx >> s) & 1) != o] = 0
            b /= np.linalg.norm(b)
        assert abs(surpr + logp) < 1e-10
        assert np.abs(a - b).max() < 1e-11


def test_sequential_sampling_statistics():
    # Outcome frequencies follow Born's rule (cusv_equiv.cpp:281-340 z-test).
    rng = np.random.default_rng(4)
    psi0 = rng.normal(size=8) + 1j * rng.normal(size=8)
    psi0 /= np.linalg.norm(psi0)
    probs = np.abs(psi0) ** 2
    counts = np.zeros(8)
    trials = 20000
    for t in range(trials):
        psi = psi0.copy()
        idx = 0
        for q in range(3):
            st, o, _ = O.measure_collapse(psi, 3, q, float(rng.random()))
            idx |= o << q
        counts[idx] += 1
    z = (counts - trials * probs) / np.sqrt(trials * probs * (1 - probs))
    assert np.abs(z).max() < 5


# ------------------------------------------------------------------ entropies

def test_entropy_primitives_kat():
    # tests/core_tests.cpp:736-800
    for d in (2, 4, 8):
        assert abs(O.entropy_from_eigs(np.full(d, 1.0 / d)) - math.log2(d)) < 1e-12
        assert abs(O.renyi2_from_purity(1.0 / d) - math.log2(d)) < 1e-12
    assert O.entropy_from_eigs([1.0, 0.0]) == 0.0
    assert O.renyi2_from_purity(1.0 + 1e-9) == 0.0
    q = 0.3
    assert abs(O.renyi2_from_purity(q * q + (1 - q) ** 2) + math.log2(q * q + (1 - q) ** 2)) < 1e-12
    assert abs(O.entropy_from_eigs([0.5, 0.5]) - 1.0) < 1e-12


@needs_ref
def test_entropy_cutoffs_match_reference_code():
    for lam in [0.0, 1e-16, 1e-15, 2e-15, 1e-3, 0.5, 1.0]:
        want = REF.ref_entropy_from_eigenvalue(lam)
        got = O.entropy_from_eigs([lam]) if lam > 1e-15 else 0.0
        assert abs(got - (0.0 if abs(want) < 1e-12 else want)) < 1e-15
    assert REF.ref_finish_entropy(5e-13) == 0.0


def random_unitary(n, rng):
    z = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


@pytest.mark.parametrize("n", [2, 4, 8, 16, 24])
def test_jacobi_known_spectrum(n):
    # tests/spectral_tests.cpp:142-187: conjugate a known spectrum with a wide
    # dynamic range and an exact degeneracy by a random unitary.
    rng = np.random.default_rng(n)
    spec = rng.normal(size=n)
    spec[0] *= 1e4
    spec[-1] = spec[-2]
    u = random_unitary(n, rng)
    a = (u * spec) @ u.conj().T
    w = np.sort(O.jacobi_eigvalsh(a))
    assert np.abs(w - np.sort(spec)).max() < 1e-9 * np.abs(spec).max()
    if REF is not None and n in (2, 4, 8, 16):
        wr = np.zeros(n)
        assert REF.ref_hermitian_eigenvalues(np.ascontiguousarray(a).ctypes.data_as(O._dp), n, O._d(wr)) == 0
        assert np.abs(np.sort(wr) - w).max() < 1e-9 * np.abs(spec).max()


def bell_pairs_state(n, pairs):
    """Product of Bell pairs (a,b) on n qubits, other qubits |0>."""
    psi = np.zeros(1 << n, np.complex128)
    for bits in itertools.product([0, 1], repeat=len(pairs)):
        idx = 0
        for (a, b), v in zip(pairs, bits):
            if v:
                idx |= (1 << a) | (1 << b)
        psi[idx] = 1.0
    return psi / np.linalg.norm(psi)


def test_half_chain_bell_pairs_closed_form():
    # Bell pairs straddling the cut contribute exactly one bit each.
    n = 8
    psi = bell_pairs_state(n, [(0, 7), (3, 4), (1, 5)])
    assert abs(O.half_chain_entropy(psi, n) - 3.0) < 1e-12
    psi = bell_pairs_state(n, [(0, 1), (4, 5)])
    assert O.half_chain_entropy(psi, n) == 0.0


def test_tmi_closed_forms():
    n = 4
    ghz = np.zeros(16, np.complex128)
    ghz[0] = ghz[15] = 1 / math.sqrt(2)
    assert abs(O.tmi(ghz, n, 1) - 1.0) < 1e-12  # 4-party GHZ: I3 = +1
    prod = O.zero_state(8)
    assert O.tmi(prod, 8) == 0.0
    # pairwise Bell pairs between regions: I3 = 0
    psi = bell_pairs_state(8, [(0, 2), (1, 5), (3, 6)])
    assert abs(O.tmi(psi, 8)) < 1e-12


def ghz_state(n, sites):
    psi = np.zeros(1 << n, np.complex128)
    psi[0] = psi[sum(1 << q for q in sites)] = 1 / math.sqrt(2)
    return psi


def i3_of(psi, n, A, B, C, D):
    s = lambda k: O.entropy(psi, n, k)
    return s(A) + s(B) + s(C) + s(D) - s(A + B) - s(A + C) - s(B + C)


def wrong_layouts(n, b, o):
    """Plausible mistakes in the cyclic regions of compute.hpp:64-113: offset
    off by one either way, the blocks laid out in the other direction, and a
    D of width b (the last n - 4b sites dropped)."""
    cyc = lambda start, k, step=1: [(start + step * i) % n for i in range(k)]
    out = []
    for oo in (o - 1, o + 1):
        out.append((cyc(oo, b), cyc(oo + b, b), cyc(oo + 2 * b, b), cyc(oo + 3 * b, n - 3 * b)))
    out.append((cyc(o, b, -1), cyc(o - b, b, -1), cyc(o - 2 * b, b, -1), cyc(o - 3 * b, n - 3 * b, -1)))
    if n > 4 * b:
        out.append((cyc(o, b), cyc(o + b, b), cyc(o + 2 * b, b), cyc(o + 3 * b, b)))
    return out


@pytest.mark.parametrize("n,b", [(8, 2), (10, 2), (12, 3), (14, 3)])
def test_tmi_regions_pinned_by_ghz4(n, b):
    # A 4-party GHZ state has I3 = +1 bit exactly when its four sites fall in
    # four different regions and 0 otherwise (S_X = 1 iff X holds some but not
    # all GHZ sites).  GHZ on (last of A, first of B, last of C, last of D)
    # gives +1 for the regions of compute.hpp:64-113 at every cycle offset,
    # and 0 under each plausible mistake in them (checked here explicitly, so
    # the pin is known to discriminate).
    for o in range(b):
        A, B, C, D = O.tmi_regions(n, b, o)
        assert A == [(o + i) % n for i in range(b)] and len(D) == n - 3 * b
        sites = [(o + b - 1) % n, (o + b) % n, (o + 3 * b - 1) % n, (o + n - 1) % n]
        psi = ghz_state(n, sites)
        assert abs(O.tmi(psi, n, b, o) - 1.0) < 1e-12
        for lay in wrong_layouts(n, b, o):
            assert abs(i3_of(psi, n, *lay)) < 1e-12, (o, lay)
        # and the offset is not interchangeable: the other cycles give 0
        for o2 in range(b):
            if o2 != o:
                assert abs(O.tmi(psi, n, b, o2)) < 1e-12


@needs_ref
def test_subsystem_entropy_matches_reference_code():
    # entropy_subsystem_rdm_from_state (manual RDM + Jacobi: no LAPACK here).
    rng = np.random.default_rng(7)
    n = 8
    psi, _, _, _ = O.simulate(n, 12, 0.1, 99)
    for kept in ([0, 1, 2], [3, 7, 1], [2, 3, 4, 5], [0], [0, 1, 2, 3, 4]):
        k = np.array(kept, np.int32)
        want = REF.ref_entropy_subsystem_rdm(psi.ctypes.data_as(O._dp), n, O._i(k), len(kept), 1)
        got = O.entropy(psi, n, kept)
        assert abs(got - want) < 1e-10
        want2 = REF.ref_entropy_subsystem_rdm(psi.ctypes.data_as(O._dp), n, O._i(k), len(kept), 2)
        assert abs(O.entropy(psi, n, kept, order=2) - want2) < 1e-10
    # the C++ oracle RDM agrees with the numpy gather
    kept = np.array([3, 7, 1], np.int32)
    rho = np.zeros((8, 8), np.complex128)
    O.lib().orc_rdm(psi.ctypes.data_as(O._dp), n, O._i(kept), 3, rho.ctypes.data_as(O._dp))
    assert np.abs(rho - O.rdm(psi, n, [3, 7, 1])).max() < 1e-14


def page_entropy_bits(m, nn):
    """Page's mean entanglement entropy S(m, n) for m <= n, in bits."""
    s = sum(1.0 / k for k in range(nn + 1, m * nn + 1)) - (m - 1) / (2 * nn)
    return s / math.log(2)


def test_page_closed_form_tmi_value():
    # cpp/README.md:456-458: deep p=0 Haar at N=16 gives I3 = -5.847228 bits;
    # it is 4*S(2^4, 2^12) - 3*S(2^8, 2^8) by Page's formula.
    # (The formula gives -5.847223; the README quotes -5.847228.)
    i3 = 4 * page_entropy_bits(16, 4096) - 3 * page_entropy_bits(256, 256)
    assert abs(i3 - (-5.847228)) < 1e-5


def test_deep_random_circuit_reaches_page_value():
    # Monte-Carlo Page check at N=8: p=0 brickwork of depth 4N is close to a
    # Haar-random state, whose half-chain entropy is Page's value.
    n = 8
    vals = []
    for s in range(1, 41):
        psi, _, _, _ = O.simulate(n, 4 * n, 0.0, s)
        vals.append(O.half_chain_entropy(psi, n))
        assert abs(np.vdot(psi, psi).real - 1) < 1e-12
    page = page_entropy_bits(16, 16)
    assert abs(np.mean(vals) - page) < 4 * np.std(vals) / math.sqrt(len(vals)) + 0.01


def test_probe_bell_pair_rho1():
    # A Bell pair of site 2 with the probe (qubit n) gives S = ln 2 nats and a
    # zero Bloch vector (cpp/include/mipt/probed.hpp:25-33, ancilla_mi.hpp:24-52).
    n = 4
    psi = np.zeros(1 << (n + 1), np.complex128)
    psi[0] = psi[(1 << 2) | (1 << n)] = 1 / math.sqrt(2)
    r = O.rho1(psi, n + 1, n)
    assert abs(O.rho1_entropy_nats(r) - math.log(2)) < 1e-14
    assert np.abs(O.bloch_from_rho1(r)).max() < 1e-14
    # a pure probe state |+i> has Bloch (0,1,0)
    psi = np.zeros(4, np.complex128)
    psi[0] = 1 / math.sqrt(2)
    psi[2] = 1j / math.sqrt(2)
    r = O.rho1(psi, 2, 1)
    assert np.abs(O.bloch_from_rho1(r) - np.array([0, 1, 0])).max() < 1e-14
    assert O.rho1_entropy_nats(r) < 1e-12


def test_measured_chain_state_stays_normalised_and_outcomes_recorded():
    n = 10
    psi, outcomes, p1, (g, f, u) = O.simulate(n, 20, 0.3, 5)
    assert abs(np.vdot(psi, psi).real - 1) < 1e-12
    assert np.all((outcomes >= 0) == (f == 1))
    m = f == 1
    assert np.all(((u[m] < p1[m]) == (outcomes[m] == 1)))
    # a measured site is in a Z eigenstate right after the layer
    psi2 = O.zero_state(n)
    O.run_layers(psi2, n, g, f, u, 0, 1)
    idx = np.arange(1 << n)
    for s in range(n):
        if f[0, s]:
            w1 = np.sum(np.abs(psi2[((idx >> s) & 1) == 1]) ** 2)
            assert min(w1, 1 - w1) < 1e-14


def test_probe_attach_closed_forms():
    # reset(site) + h(site) + cx(site, probe) on a scrambled state: the site is
    # disentangled by the reset, so {site, probe} is a pure Bell pair:
    # S(probe) = ln 2 (nats), S({site, probe}) = 0, and the reset outcome
    # follows Born's rule for the site.
    rng = np.random.default_rng(5)
    n, site = 6, 2
    for u in (0.05, 0.5, 0.95):
        phi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        phi /= np.linalg.norm(phi)
        psi = np.zeros(1 << (n + 1), np.complex128)
        psi[: 1 << n] = phi
        p1 = float(np.sum(np.abs(phi[(np.arange(1 << n) >> site) & 1 == 1]) ** 2))
        o = O.attach_probe(psi, n, site, u)
        assert o == (1 if u < p1 else 0)
        assert abs(np.linalg.norm(psi) - 1.0) < 1e-12
        r = O.rho1(psi, n + 1, n)
        assert abs(O.rho1_entropy_nats(r) - np.log(2)) < 1e-12
        assert abs(O.entropy(psi, n + 1, [site, n])) < 1e-10
        # the other qubits hold phi collapsed on site = o
        keep = np.arange(1 << n)
        col = phi[(keep >> site) & 1 == o]
        rest = psi[(np.arange(1 << (n + 1)) & ((1 << site) | (1 << n))) == 0] * np.sqrt(2.0)
        assert np.allclose(np.abs(rest), np.abs(col) / np.linalg.norm(col), atol=1e-12)


def test_probe_protocol_oracle_runs_and_purifies():
    # p = 1 after the attach measures every site each layer: the probe purifies
    psi, outs, p1, att, _ = O.probe_trajectory(8, 12, 1.0, 77, 4, 4)
    assert att in (0, 1)
    assert np.all(outs[:4] == -1) and np.all(outs[4:] >= 0)
    r = O.rho1(psi, 9, 8)
    assert O.rho1_entropy_nats(r) < 1e-9
    # p = 0: the probe stays maximally entangled
    psi, *_ = O.probe_trajectory(8, 12, 0.0, 78, 4, 4)
    assert abs(O.rho1_entropy_nats(O.rho1(psi, 9, 8)) - np.log(2)) < 1e-10


@pytest.mark.parametrize("gx,gy,closed", [(5, 4, False), (4, 5, True), (4, 3, False), (2, 3, True)])
def test_grid_bond_layout(gx, gy, closed):
    # four colours: each site in at most one bond per layer, the union is
    # every nearest-neighbour pair of the grid (rows closed when asked and gx
    # is even), and the engine's host restatement lists the same bonds in the
    # same (gate-stream) order
    import mipt_amd
    n = gx * gy
    seen = set()
    for t in range(4):
        bonds = O.grid_layer_bonds(gx, gy, t, closed)
        sites = [q for b in bonds for q in b]
        assert len(sites) == len(set(sites))
        assert bonds == mipt_amd.layer_bonds(n, t, closed, gx)
        assert bonds == O.grid_layer_bonds(gx, gy, t + 4, closed)
        seen |= {tuple(sorted(b)) for b in bonds}
    want = {(r * gx + c, r * gx + c + 1) for r in range(gy) for c in range(gx - 1)}
    want |= {(r * gx + c, (r + 1) * gx + c) for r in range(gy - 1) for c in range(gx)}
    if closed and gx % 2 == 0 and gx > 2:
        want |= {(r * gx, r * gx + gx - 1) for r in range(gy)}
    assert seen == want


def test_single_gate_stream_matches_layer_builder():
    g, f, u = O.build_layers(8, 3, 0.2, 5)
    i = 0
    x = np.zeros(32)
    for t in range(3):
        for b, _ in enumerate(O.layer_bonds(8, t, True)):
            O.lib().orc_philox_haar(5, t, b, x.ctypes.data_as(O._dp))
            assert np.array_equal(x[0::2] + 1j * x[1::2], g[i].ravel())
            i += 1
