"""GPU: several reference qubits through the batched engine.

Trajectories with 2, 3 and 4 references (Bell pairs at t = 0, or attached
after unmeasured layers: reset + h + cx per reference, in order) against the
FP64 oracle: outcomes, states (1e-5), the references' joint density matrix
(1e-5 per entry) and its information measures (1e-4 nats).  Closed form on
the device: a Bell measurement of the two sites the references are paired
with leaves the references in a Bell state (S(AB) = 0, I(A:B) = 2 ln 2,
negativity 1/2), in the one-block kernel (L=4) and the tiled sweeps (L=16).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_engine import check_outcomes
from tests.test_references import swap_layers

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    import mipt_amd
    mipt_amd.lib()
    return mipt_amd


CASES = [  # L, layers, p, seed, reference sites, unmeasured layers before the attach
    (8, 10, 0.15, 41, (2, 5), 0),
    (8, 12, 0.2, 42, (1, 4, 6), 4),
    (10, 10, 0.15, 43, (1, 3, 6, 9), 3),
    (12, 10, 0.15, 44, (2, 5, 8, 11), 4),
    (16, 8, 0.1, 45, (7, 15), 0),
    (16, 10, 0.15, 46, (3, 7, 11, 15), 3),
    (18, 8, 0.1, 47, (0, 9, 17), 2),
]
# references attached after monitored layers (the two- and four-reference
# protocols attach at t0 = 2N after measured evolution): attach layer > the
# unmeasured prefix, so the previous layer's measurements are applied first
ATTACH_CASES = [  # L, layers, p, seed, sites, unmeasured prefix, attach layer
    (8, 12, 0.2, 51, (1, 5), 0, 4),
    (10, 12, 0.2, 52, (0, 2, 5, 9), 1, 5),
    (14, 10, 0.2, 53, (3, 10), 0, 4),
    (16, 10, 0.15, 54, (3, 7, 11, 15), 2, 5),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_batch_references_match_oracle(M, case):
    L, layers, p, seed, sites, pz = CASES[case]
    k = len(sites)
    print(f"references L={L} k={k} p_zero_layers={pz}", flush=True)
    b = M.Batch([M.Spec(L, layers, p, seed, True, probe_sites=sites, p_zero_layers=pz)])
    b.generate()
    b.evolve()
    b.synchronize()
    psi, outs, p1, att, (g, f, u) = O.references_trajectory(L, layers, p, seed, list(sites), pz)
    got, flagged, _, _ = b.outcomes(0)
    check_outcomes(got, outs, p1, u, f, f"case {case}")
    rho = b.probe_rdm()[0]
    assert abs(np.trace(rho) - 1) < 1e-12 and np.abs(rho - rho.conj().T).max() < 1e-15
    if flagged.any():
        pytest.skip("a draw within the flag margin of P(1): trajectories may differ")
    assert np.abs(b.get_state(0) - psi).max() < 1e-5
    want = O.reference_rdm(psi, L, k)
    assert np.abs(rho - want).max() < 1e-5
    gi, wi = M.probe_information(rho), O.probe_information(want)
    for key in wi:
        assert abs(gi[key] - wi[key]) < 1e-4, (key, gi[key], wi[key])
    b.close()


@pytest.mark.parametrize("case", range(len(ATTACH_CASES)))
def test_batch_references_attached_after_monitored_layers(M, case):
    L, layers, p, seed, sites, pz, att_layer = ATTACH_CASES[case]
    k = len(sites)
    print(f"references L={L} k={k} p_zero_layers={pz} attach_layer={att_layer}", flush=True)
    b = M.Batch([M.Spec(L, layers, p, seed, True, probe_sites=sites, p_zero_layers=pz, attach_layer=att_layer)])
    b.generate()
    b.evolve()
    b.synchronize()
    psi, outs, p1, att, (g, f, u) = O.references_trajectory(L, layers, p, seed, list(sites), pz,
                                                            attach_layer=att_layer)
    assert f[:att_layer].sum() > 0  # measurements before the attach
    got, flagged, _, _ = b.outcomes(0)
    check_outcomes(got, outs, p1, u, f, f"attach case {case}")
    if flagged.any():
        pytest.skip("a draw within the flag margin of P(1): trajectories may differ")
    assert np.abs(b.get_state(0) - psi).max() < 1e-5
    rho, want = b.probe_rdm()[0], O.reference_rdm(psi, L, k)
    assert np.abs(rho - want).max() < 1e-5
    gi, wi = M.probe_information(rho), O.probe_information(want)
    for key in wi:
        assert abs(gi[key] - wi[key]) < 1e-4, (key, gi[key], wi[key])
    b.close()


STAGGER_CASES = [  # L, layers, p, seed, sites, attach layer, stagger (a site may serve both)
    (8, 14, 0.15, 61, (2, 2), 4, 3),
    (10, 12, 0.2, 62, (1, 6), 2, 5),
    (16, 12, 0.15, 63, (5, 5), 3, 4),
]


@pytest.mark.parametrize("case", range(len(STAGGER_CASES)))
def test_batch_references_staggered_attach(M, case):
    L, layers, p, seed, sites, att_layer, stagger = STAGGER_CASES[case]
    print(f"staggered references L={L} sites={sites} attach={att_layer}+{stagger}j", flush=True)
    b = M.Batch([M.Spec(L, layers, p, seed, True, probe_sites=sites, attach_layer=att_layer,
                        attach_stagger=stagger)])
    b.generate()
    b.evolve()
    b.synchronize()
    psi, outs, p1, att, (g, f, u) = O.references_trajectory(L, layers, p, seed, list(sites), 0,
                                                            attach_layer=att_layer, attach_stagger=stagger)
    assert len(att) == 2
    got, flagged, _, _ = b.outcomes(0)
    check_outcomes(got, outs, p1, u, f, f"stagger case {case}")
    if flagged.any():
        pytest.skip("a draw within the flag margin of P(1): trajectories may differ")
    assert np.abs(b.get_state(0) - psi).max() < 1e-5
    rho, want = b.probe_rdm()[0], O.reference_rdm(psi, L, 2)
    assert np.abs(rho - want).max() < 1e-5
    b.close()


@pytest.mark.parametrize("n,s0", [(4, 1), (16, 7)])
def test_entanglement_swapping_on_device(M, n, s0):
    print(f"entanglement swapping L={n}", flush=True)
    b = M.Batch([M.Spec(n, 2, 0.0, 1, True, probe_sites=(s0, s0 + 1))])
    g, f, u = swap_layers(n, s0)
    b.set_layers(0, g, f, u)
    b.evolve()
    b.synchronize()
    info = M.probe_information(b.probe_rdm()[0])
    assert abs(info["sq"]) < 1e-5
    assert abs(info["i2"] - 2 * math.log(2)) < 1e-5
    assert abs(info["negativity"] - 0.5) < 1e-5
    b.close()


def test_reference_spec_validation(M):
    bad = [M.Spec(8, 4, 0.1, 1, True, probe_sites=(2, 2)),            # repeated site
           M.Spec(8, 4, 0.1, 1, True, probe_site=3, probe_sites=(2, 5)),  # both forms
           M.Spec(8, 4, 0.1, 1, True, probe_sites=(2, 8)),            # site outside the chain
           M.Spec(24, 4, 0.1, 1, True, probe_sites=(1, 2, 3, 4)),     # 28 qubits
           M.Spec(8, 4, 0.1, 1, True, probe_sites=(1, 2), attach_layer=4)]  # attach past the last layer
    for s in bad:
        with pytest.raises(M.MiptHipError):
            M.Batch([s])
    b = M.Batch([M.Spec(8, 4, 0.1, 1, True, probe_sites=(1, 2)), M.Spec(8, 4, 0.1, 2, True, probe_sites=(1, 2, 3))])
    b.generate()
    b.evolve()
    b.synchronize()
    with pytest.raises(M.MiptHipError):
        b.probe_rdm()  # mixed reference counts


@pytest.mark.parametrize("L", [8, 16])
def test_batch_mixed_attach_layers_in_one_group(M, L):
    # trajectories of one sweep group with different attach layers: every
    # attach layer applies the pending measurements of the whole group first
    specs = [M.Spec(L, 10, 0.2, 71, True, probe_sites=(3, L - 2), attach_layer=2),
             M.Spec(L, 10, 0.2, 72, True, probe_sites=(1, 5), attach_layer=4),
             M.Spec(L, 10, 0.2, 73, True, probe_sites=(6, 6), attach_layer=2, attach_stagger=3),
             M.Spec(L, 10, 0.2, 74, True, probe_sites=(0, 4), p_zero_layers=3)]
    print(f"mixed attach layers L={L}", flush=True)
    b = M.Batch(specs)
    b.generate()
    b.evolve()
    b.synchronize()
    rho = b.probe_rdm()
    for i, s in enumerate(specs):
        psi, outs, p1, att, (g, f, u) = O.references_trajectory(
            L, s.layers, s.p, s.seed, list(s.probe_sites), s.p_zero_layers,
            attach_layer=None if s.attach_layer < 0 else s.attach_layer, attach_stagger=s.attach_stagger)
        got, flagged, _, _ = b.outcomes(i)
        check_outcomes(got, outs, p1, u, f, f"mixed traj {i}")
        if flagged.any():
            continue
        assert np.abs(b.get_state(i) - psi).max() < 1e-5, i
        assert np.abs(rho[i] - O.reference_rdm(psi, L, 2)).max() < 1e-5, i
    b.close()
